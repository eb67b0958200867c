// C ABI of the metriko HIP library: contexts, operators, the field stage and helpers.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <numeric>
#include "mk_internal.h"
#include "sell.h"

using mk::DevBuf;
using mk::Error;

namespace {

template <typename Fn>
mk_status guard(mk_ctx* c, Fn&& fn) {
    try {
        fn();
        return MK_OK;
    } catch (const Error& e) {
        if (c) c->last_error = e.what();
        return e.code;
    } catch (const std::bad_alloc&) {
        if (c) c->last_error = "host allocation failed";
        return MK_ERR_ALLOC;
    } catch (const std::exception& e) {
        if (c) c->last_error = e.what();
        return MK_ERR_ARG;
    }
}

double now_ms() {
    return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

mk_solve_opts resolve(const mk_solve_opts* o) {
    mk_solve_opts d;
    mk_default_opts(&d);
    return o ? *o : d;
}

// Connection Laplacian rows (face_rosy_field.h:46-76), one thread per face, written
// directly into a SELL layout of width 4 (diagonal + up to three neighbours, columns
// sorted; missing neighbours on boundary faces are zero padding on the diagonal column).
// transport across twin(h) into face f: r = -e^{i (farg[h] - farg[twin h])}; the
// off-diagonal entry is -r^N = -(-1)^N e^{i N (farg[h] - farg[twin h])}.
__global__ void k_field_laplacian(int nF, int C, const int* __restrict__ fh, const int* __restrict__ twin,
                                  const int* __restrict__ hface, const double* __restrict__ farg, int N,
                                  int* __restrict__ col, double* __restrict__ val, int* __restrict__ cnt) {
    const int f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= nF) return;
    int c[4];
    double vr[4], vi[4];
    int m = 0;
    double diag = 0.0;
    const double sgn = (N % 2 == 0) ? -1.0 : 1.0;
    for (int k = 0; k < 3; k++) {
        const int h = fh[3 * f + k];
        const int tw = twin[h];
        const int g = hface[tw];
        if (g < 0) continue;
        diag += 1.0;
        double s, co;
        sincos(double(N) * (farg[h] - farg[tw]), &s, &co);
        c[m] = g;
        vr[m] = sgn * co;
        vi[m] = sgn * s;
        m++;
    }
    c[m] = f;
    vr[m] = diag + 1e-9;
    vi[m] = 0.0;
    m++;
    // insertion sort by column (m <= 4)
    for (int i = 1; i < m; i++) {
        int cc = c[i];
        double a = vr[i], b = vi[i];
        int j = i - 1;
        while (j >= 0 && c[j] > cc) { c[j + 1] = c[j]; vr[j + 1] = vr[j]; vi[j + 1] = vi[j]; j--; }
        c[j + 1] = cc; vr[j + 1] = a; vi[j + 1] = b;
    }
    const int s = f / C, lane = f - s * C;
    const int64_t base = int64_t(s) * 4 * C + lane;
    for (int j = 0; j < 4; j++) {
        const int64_t e = base + int64_t(j) * C;
        if (j < m) { col[e] = c[j]; val[2 * e] = vr[j]; val[2 * e + 1] = vi[j]; }
        else { col[e] = f; val[2 * e] = 0.0; val[2 * e + 1] = 0.0; }
    }
    cnt[f] = m;
}

__global__ void k_to_f32(size_t len, const double* x, float* y) {
    const size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < len) y[i] = float(x[i]);
}
__global__ void k_fill_i32(size_t len, int v, int* y) {
    const size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < len) y[i] = v;
}
__global__ void k_slice_ptr(int ns, int64_t w, int64_t C, int64_t* ptr, int* sw) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s <= ns) ptr[s] = int64_t(s) * w * C;
    if (s < ns) sw[s] = int(w);
}
__global__ void k_dinv_from_diag(int n, int C, const int* col, const double* val, double* d64, float* d32) {
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= n) return;
    const int s = r / C, lane = r - s * C;
    double d = 0.0;
    for (int j = 0; j < 4; j++) {
        const int64_t e = int64_t(s) * 4 * C + int64_t(j) * C + lane;
        if (col[e] == r) d += val[2 * e];
    }
    d64[r] = d != 0.0 ? 1.0 / d : 1.0;
    d32[r] = float(d64[r]);
}

inline unsigned nblk(size_t n, int bs = 256) { return unsigned((n + bs - 1) / bs); }

// ---- Galerkin start of the inverse iteration -------------------------------------------------
// The window holds the last kWin iterates u (complex, 2n doubles each, column j at W + 2n j).
constexpr int kWin = 4;
constexpr int kWinDots = 4 * kWin;  // per column: w^H a (re, im), w^H b (re, im)

// Per-block partial sums of w_j^H a and w_j^H b for the nw <= kWin window columns.
__global__ void k_window_dots(int n, int nw, const double* __restrict__ W, const double* __restrict__ a,
                              const double* __restrict__ b, double* __restrict__ part) {
    double acc[kWinDots];
#pragma unroll
    for (int k = 0; k < kWinDots; k++) acc[k] = 0.0;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const double ar = a[2 * i], ai = a[2 * i + 1], br = b[2 * i], bi = b[2 * i + 1];
#pragma unroll
        for (int j = 0; j < kWin; j++) {
            if (j < nw) {
                const double wr = W[size_t(j) * 2 * n + 2 * i], wi = W[size_t(j) * 2 * n + 2 * i + 1];
                acc[4 * j] += wr * ar + wi * ai;
                acc[4 * j + 1] += wr * ai - wi * ar;
                acc[4 * j + 2] += wr * br + wi * bi;
                acc[4 * j + 3] += wr * bi - wi * br;
            }
        }
    }
    mk::block_sum<kWinDots>(acc);
    if (threadIdx.x == 0)
        for (int k = 0; k < kWinDots; k++) part[size_t(blockIdx.x) * kWinDots + k] = acc[k];
}
// Fixed-order sum of the per-block partials: one wavefront per value, lane l adds the
// partials l, l + 64, ... and the wavefront's fixed shuffle tree combines the lanes.
__global__ void __launch_bounds__(64 * kWinDots) k_window_sum(int nparts, const double* __restrict__ part,
                                                               double* __restrict__ out) {
    const int k = threadIdx.x >> 6, lane = threadIdx.x & 63;
    double s = 0.0;
    for (int p = lane; p < nparts; p += 64) s += part[size_t(p) * kWinDots + k];
    s = mk::wave_sum(s);
    if (lane == 0) out[k] = s;
}
// x = sum_j y_j w_j (complex coefficients y[2j], y[2j+1]).
__global__ void k_window_combine(int n, int nw, const double* __restrict__ W, const double* __restrict__ y,
                                 double* __restrict__ x) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double xr = 0.0, xi = 0.0;
    for (int j = 0; j < nw; j++) {
        const double wr = W[size_t(j) * 2 * n + 2 * i], wi = W[size_t(j) * 2 * n + 2 * i + 1];
        xr += y[2 * j] * wr - y[2 * j + 1] * wi;
        xi += y[2 * j] * wi + y[2 * j + 1] * wr;
    }
    x[2 * i] = xr;
    x[2 * i + 1] = xi;
}

}  // namespace

namespace mk {
// Eigen-decomposition of a real symmetric m x m matrix (cyclic Jacobi); A is destroyed, V
// receives the eigenvectors as columns, w the eigenvalues.
void jacobi_eig(int m, std::vector<double>& A, std::vector<double>& V, std::vector<double>& w) {
    V.assign(size_t(m) * m, 0.0);
    for (int i = 0; i < m; i++) V[size_t(i) * m + i] = 1.0;
    for (int sweep = 0; sweep < 60; sweep++) {
        double off = 0.0, tot = 0.0;
        for (int p = 0; p < m; p++)
            for (int q = 0; q < m; q++) {
                const double v = A[size_t(p) * m + q] * A[size_t(p) * m + q];
                tot += v;
                if (p != q) off += v;
            }
        if (off <= 1e-30 * tot) break;
        for (int p = 0; p < m; p++)
            for (int q = p + 1; q < m; q++) {
                const double apq = A[size_t(p) * m + q];
                if (std::abs(apq) < 1e-300) continue;
                const double th = (A[size_t(q) * m + q] - A[size_t(p) * m + p]) / (2.0 * apq);
                const double t = (th >= 0 ? 1.0 : -1.0) / (std::abs(th) + std::sqrt(th * th + 1.0));
                const double cs = 1.0 / std::sqrt(t * t + 1.0), sn = t * cs;
                for (int k = 0; k < m; k++) {  // A P
                    const double akp = A[size_t(k) * m + p], akq = A[size_t(k) * m + q];
                    A[size_t(k) * m + p] = cs * akp - sn * akq;
                    A[size_t(k) * m + q] = sn * akp + cs * akq;
                }
                for (int k = 0; k < m; k++) {  // P^T (A P)
                    const double apk = A[size_t(p) * m + k], aqk = A[size_t(q) * m + k];
                    A[size_t(p) * m + k] = cs * apk - sn * aqk;
                    A[size_t(q) * m + k] = sn * apk + cs * aqk;
                }
                for (int k = 0; k < m; k++) {  // V P
                    const double vkp = V[size_t(k) * m + p], vkq = V[size_t(k) * m + q];
                    V[size_t(k) * m + p] = cs * vkp - sn * vkq;
                    V[size_t(k) * m + q] = sn * vkp + cs * vkq;
                }
            }
    }
    w.resize(size_t(m));
    for (int i = 0; i < m; i++) w[size_t(i)] = A[size_t(i) * m + i];
}

// y = G^+ g for the Hermitian positive semi-definite G (m x m, row-major), through its real
// embedding [[Re G, -Im G], [Im G, Re G]]; eigenvalues below 1e-13 of the largest are
// dropped (the window's columns become nearly parallel as the iteration converges).
std::vector<std::complex<double>> herm_pinv_solve(int m, const std::vector<std::complex<double>>& G,
                                                  const std::vector<std::complex<double>>& g) {
    const int r = 2 * m;
    std::vector<double> A(static_cast<size_t>(r) * r), V, w, rhs(static_cast<size_t>(r));
    for (int i = 0; i < m; i++) {
        for (int j = 0; j < m; j++) {
            const std::complex<double> z = G[size_t(i) * m + j];
            A[size_t(i) * r + j] = z.real();
            A[size_t(i) * r + j + m] = -z.imag();
            A[size_t(i + m) * r + j] = z.imag();
            A[size_t(i + m) * r + j + m] = z.real();
        }
        rhs[size_t(i)] = g[size_t(i)].real();
        rhs[size_t(i + m)] = g[size_t(i)].imag();
    }
    jacobi_eig(r, A, V, w);
    double wmax = 0.0;
    for (double v : w) wmax = std::max(wmax, v);
    std::vector<double> y(static_cast<size_t>(r), 0.0);
    for (int k = 0; k < r; k++) {
        if (!(w[size_t(k)] > 1e-13 * wmax)) continue;
        double c = 0.0;
        for (int i = 0; i < r; i++) c += V[size_t(i) * r + k] * rhs[size_t(i)];
        c /= w[size_t(k)];
        for (int i = 0; i < r; i++) y[size_t(i)] += c * V[size_t(i) * r + k];
    }
    std::vector<std::complex<double>> out(static_cast<size_t>(m));
    for (int i = 0; i < m; i++) out[size_t(i)] = {y[size_t(i)], y[size_t(i + m)]};
    return out;
}

// host CSR copy of an operator (from its host arrays, or compacted from the device SELL layout)
void host_csr(mk_ctx* c, const mk_op* A, std::vector<int64_t>& rp, std::vector<int>& cc, std::vector<double>& vv) {
    const int vw = A->cplx ? 2 : 1;
    if (!A->h_ptr.empty()) {
        rp.assign(A->h_ptr.begin(), A->h_ptr.end());
        cc.assign(A->h_col.begin(), A->h_col.end());
        vv.assign(A->h_val.begin(), A->h_val.end());
        return;
    }
    MK_CHECK(A->nmid == 0, MK_ERR_UNSUPPORTED, "operator export: no host copy of the medium-row part");
    std::vector<int64_t> sptr(size_t(A->nslices) + 1);
    std::vector<int> sw(size_t(A->nslices)), hc(size_t(A->stored));
    std::vector<double> hv(size_t(A->stored) * vw);
    A->slice_ptr.download(sptr.data(), sptr.size(), c->stream);
    A->slice_w.download(sw.data(), sw.size(), c->stream);
    A->col.download(hc.data(), hc.size(), c->stream);
    A->val64.download(hv.data(), hv.size(), c->stream);
    std::vector<int> perm;
    if (A->perm.p) {
        perm.resize(size_t(A->n_short));
        A->perm.download(perm.data(), perm.size(), c->stream);
    }
    std::vector<int> lrows(size_t(A->nlong)), lcol;
    std::vector<int64_t> lptr(size_t(A->nlong) + 1, 0);
    std::vector<double> lval;
    if (A->nlong) {
        A->long_rows.download(lrows.data(), lrows.size(), c->stream);
        A->long_ptr.download(lptr.data(), lptr.size(), c->stream);
        MK_HIP(hipStreamSynchronize(c->stream));
        lcol.resize(size_t(lptr.back()));
        lval.resize(size_t(lptr.back()) * vw);
        A->long_col.download(lcol.data(), lcol.size(), c->stream);
        A->long_val64.download(lval.data(), lval.size(), c->stream);
    }
    MK_HIP(hipStreamSynchronize(c->stream));
    // row -> (kind, position): kind 0 = SELL position, 1 = long-row index
    std::vector<int> kind(size_t(A->n), -1), pos(size_t(A->n), -1);
    for (int r = 0; r < A->n_short; r++) { const int row = perm.empty() ? r : perm[r]; kind[row] = 0; pos[row] = r; }
    for (int w = 0; w < A->nlong; w++) { kind[lrows[w]] = 1; pos[lrows[w]] = w; }
    // two passes over the rows (count the stored non-zeros, then fill), OpenMP-parallel
    auto row_entries = [&](int row, auto&& emit) {
        if (kind[row] == 0) {
            const int r = pos[row];
            const int s = r / A->C, lane = r % A->C;
            for (int j = 0; j < sw[s]; j++) {
                const int64_t e = sptr[s] + int64_t(j) * A->C + lane;
                bool zero = true;
                for (int q = 0; q < vw; q++) if (hv[e * vw + q] != 0.0) zero = false;
                if (!zero) emit(hc[e], &hv[e * vw]);
            }
        } else if (kind[row] == 1) {
            const int w = pos[row];
            for (int64_t k = lptr[w]; k < lptr[w + 1]; k++) emit(lcol[k], &lval[k * vw]);
        }
    };
    rp.assign(size_t(A->n) + 1, 0);
#pragma omp parallel for schedule(static, 4096)
    for (int row = 0; row < A->n; row++) {
        int64_t cnt = 0;
        row_entries(row, [&](int, const double*) { cnt++; });
        rp[size_t(row) + 1] = cnt;
    }
    for (int row = 0; row < A->n; row++) rp[size_t(row) + 1] += rp[size_t(row)];
    cc.resize(size_t(rp[size_t(A->n)]));
    vv.resize(size_t(rp[size_t(A->n)]) * vw);
#pragma omp parallel for schedule(static, 4096)
    for (int row = 0; row < A->n; row++) {
        int64_t w = rp[size_t(row)];
        row_entries(row, [&](int col, const double* v) {
            cc[size_t(w)] = col;
            for (int q = 0; q < vw; q++) vv[size_t(w) * vw + q] = v[q];
            w++;
        });
    }
}

void ensure_mg(mk_ctx* c, const mk_op* A_, const mk_solve_opts& o) {
    mk_op* A = const_cast<mk_op*>(A_);
    if (A->mg || o.precond != MK_PRECOND_AMG) return;
    MK_CHECK(A->n == A->m, MK_ERR_ARG, "multigrid needs a square operator");
    std::vector<int64_t> rp;
    std::vector<int> cc;
    std::vector<double> vv;
    const int64_t* pp = A->h_ptr.data();
    const int* pc = A->h_col.data();
    const double* pv = A->h_val.data();
    if (A->h_ptr.empty() && !A->mg_ptr.empty()) {  // device-assembled with a host copy for the set-up
        pp = A->mg_ptr.data();
        pc = A->mg_col.data();
        pv = A->mg_val.data();
    } else if (A->h_ptr.empty()) {  // device-assembled operator: export it
        const double t0 = now_ms();
        host_csr(c, A, rp, cc, vv);
        MK_LOG("[mg] host CSR of the operator %.0f ms\n", now_ms() - t0);
        pp = rp.data();
        pc = cc.data();
        pv = vv.data();
    }
    A->mg = A->cplx ? mg_build_complex(c, A->n, pp, pc, pv, o)
                    : mg_build_real(c, A->n, pp, pc, pv, o, A->nnull ? &A->nullvec : nullptr, A->nnull);
    mg_set_labels(A->mg, A->label);
    std::vector<int64_t>().swap(A->mg_ptr);
    metriko::HVec<int>().swap(A->mg_col);
    metriko::HVec<double>().swap(A->mg_val);
}

void stats_reset(mk_ctx* c) {
    c->spmv_ms = 0.0;
    c->spmv_calls = 0;
    c->spmv_bytes = 0.0;
    c->ev_used = 0;
    c->ev_key.clear();
}
void stats_fill(mk_ctx* c, mk_stats* s) {
    timing_flush(c);
    s->spmv_ms = c->spmv_ms;
    s->spmv_calls = c->spmv_calls;
    s->spmv_bytes = c->spmv_bytes;
}
}  // namespace mk

namespace mk {
// A real operator that takes over a host CSR (no copy of its arrays); same layout as
// mk_op_create_real.
mk_op* op_create_owned(mk_ctx* c, int n, metriko::HVec<int64_t>&& ptr, metriko::HVec<int>&& col,
                       metriko::HVec<double>&& val, const mk_solve_opts& o) {
    MK_CHECK(n > 0 && ptr.size() == size_t(n) + 1 && col.size() == size_t(ptr[size_t(n)]) && val.size() == col.size(),
             MK_ERR_ARG, "operator arrays");
    MK_HIP(hipSetDevice(c->device));
    auto* op = new mk_op();
    try {
        op->ctx = c;
        op->n = n;
        op->m = n;
        op->cplx = false;
        op->nnz = ptr[size_t(n)];
        op->block = o.block_size;
        MK_CHECK(op->block % 64 == 0 && op->block >= 64 && op->block <= 256, MK_ERR_ARG,
                 "block_size must be 64, 128 or 256 (the SpMV kernels are built for at most 256 threads)");
        op->h_ptr = std::move(ptr);
        op->h_col = std::move(col);
        op->h_val = std::move(val);
        const int C = o.format == MK_FMT_CSR ? 64 : o.slice_height;
        build_sell(op, op->h_ptr.data(), op->h_col.data(), op->h_val.data(), C, o.format == MK_FMT_CSR ? 1 : o.sigma);
    } catch (...) {
        delete op;
        throw;
    }
    return op;
}
}  // namespace mk

extern "C" {

const char* mk_status_str(mk_status s) {
    switch (s) {
        case MK_OK: return "ok";
        case MK_ERR_ARG: return "invalid argument";
        case MK_ERR_HIP: return "HIP runtime error";
        case MK_ERR_NOT_SPD: return "matrix not SPD/HPD (CG breakdown)";
        case MK_ERR_NO_CONVERGENCE: return "no convergence within maxiter";
        case MK_ERR_NAN: return "NaN/Inf encountered";
        case MK_ERR_STAGNATION: return "iterative refinement stagnated";
        case MK_ERR_ALLOC: return "allocation failed";
        case MK_ERR_UNSUPPORTED: return "unsupported option";
        case MK_ERR_SINGULAR: return "singular system";
        default: return "unknown status";
    }
}

void mk_default_opts(mk_solve_opts* o) {
    o->precision = MK_PREC_MIXED_MG;
    o->tol = 1e-10;
    o->maxiter = 20000;
    o->format = MK_FMT_SELL;
    o->slice_height = 64;
    o->sigma = 1;
    o->block_size = 256;
    o->precond = MK_PRECOND_AMG;
    o->inner_tol = 1e-5;
    o->max_refine = 12;
    o->check_every = 16;
}

mk_status mk_init(mk_ctx** out, int device) {
    if (!out) return MK_ERR_ARG;
    mk_ctx* c = new (std::nothrow) mk_ctx();
    if (!c) return MK_ERR_ALLOC;
    mk_status s = guard(c, [&] {
        int n = 0;
        MK_HIP(hipGetDeviceCount(&n));
        MK_CHECK(device >= 0 && device < n, MK_ERR_ARG, "no such device");
        c->device = device;
        if (const char* g = std::getenv("METRIKO_GRAPHS")) c->use_graphs = std::atoi(g) != 0;
        MK_HIP(hipSetDevice(device));
        // the allocation-free solve's scratch on this device (a failure is reported by that
        // entry point, not here)
        (void)mk::noalloc_setup(device, nullptr);
        MK_HIP(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
        // per-block partial sums of the two-launch grid reductions (up to kMaxK quantities of
        // kMaxParts blocks each)
        c->partials.alloc(mk::kMaxParts * mk::kMaxK);
        c->state.alloc(1);
        c->scal.alloc(64);
        MK_HIP(hipMemsetAsync(c->scal.p, 0, 64 * sizeof(double), c->stream));
        MK_HIP(hipStreamSynchronize(c->stream));
        const char* t = std::getenv("METRIKO_TIME_SPMV");
        c->time_spmv = t && t[0] == '1';
    });
    if (s != MK_OK) { delete c; return s; }
    *out = c;
    return MK_OK;
}

mk_status mk_finalize(mk_ctx* c) {
    if (!c) return MK_ERR_ARG;
    (void)hipSetDevice(c->device);
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    for (auto e : c->ev_pool) (void)hipEventDestroy(e);
    c->ev_pool.clear();
    if (c->igm_state) mk_igm_state_destroy(c->igm_state);
    c->igm_state = nullptr;
    c->work = mk::Work();
    c->partials.release();
    c->state.release();
    c->scal.release();
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
    return MK_OK;
}

mk_status mk_last_error(mk_ctx* c, char* buf, int len) {
    if (!c || !buf || len <= 0) return MK_ERR_ARG;
    std::strncpy(buf, c->last_error.c_str(), size_t(len - 1));
    buf[len - 1] = 0;
    return MK_OK;
}

mk_status mk_set_timing(mk_ctx* c, int on) {
    if (!c) return MK_ERR_ARG;
    c->time_spmv = on != 0;
    return MK_OK;
}

mk_status mk_set_graphs(mk_ctx* c, int on) {
    if (!c) return MK_ERR_ARG;
    c->use_graphs = on != 0;
    return MK_OK;
}

mk_status mk_timing_report(mk_ctx* c, int reset, char* buf, int len) {
    if (!c || !buf || len <= 0) return MK_ERR_ARG;
    return guard(c, [&] {
        mk::timing_flush(c);
        std::string s = "[";
        char tmp[512];
        for (size_t i = 0; i < c->kstats.size(); i++) {
            const auto& k = c->kstats[i].second;
            std::snprintf(tmp, sizeof(tmp),
                          "%s{\"key\":\"%s\",\"calls\":%lld,\"ms\":%.6f,\"bytes\":%.1f,\"bytes_per_call\":%.1f,\"rows\":%d}",
                          i ? "," : "", c->kstats[i].first.c_str(), (long long)k.calls, k.ms, k.bytes,
                          k.bytes_per_call, k.rows);
            s += tmp;
        }
        s += "]";
        MK_CHECK(int(s.size()) < len, MK_ERR_ARG, "timing report buffer too small");
        std::memcpy(buf, s.c_str(), s.size() + 1);
        if (reset) c->kstats.clear();
    });
}

static mk_status op_create(mk_ctx* c, int n, const int64_t* rowptr, const int32_t* col, const double* val,
                           bool cplx, const mk_solve_opts* o_, mk_op** out) {
    if (!c || !out || n <= 0 || !rowptr || !col || !val) return MK_ERR_ARG;
    return guard(c, [&] {
        const mk_solve_opts o = resolve(o_);
        MK_HIP(hipSetDevice(c->device));
        auto* op = new mk_op();
        try {
            op->ctx = c;
            op->n = n;
            op->m = n;
            op->cplx = cplx;
            op->nnz = rowptr[n];
            op->block = o.block_size;
            MK_CHECK(op->block % 64 == 0 && op->block >= 64 && op->block <= 256, MK_ERR_ARG,
                 "block_size must be 64, 128 or 256 (the SpMV kernels are built for at most 256 threads)");
            for (int i = 0; i < n; i++) MK_CHECK(rowptr[i + 1] >= rowptr[i], MK_ERR_ARG, "rowptr not monotone");
            int64_t bad = 0;
#pragma omp parallel for reduction(+ : bad) schedule(static, 65536) if (op->nnz > 1000000)
            for (int64_t k = 0; k < op->nnz; k++) bad += (col[k] >= 0 && col[k] < n) ? 0 : 1;
            MK_CHECK(bad == 0, MK_ERR_ARG, "column index out of range");
            const int vw = cplx ? 2 : 1;
            op->h_ptr.assign(rowptr, rowptr + n + 1);
            op->h_col.assign(col, col + op->nnz);
            op->h_val.assign(val, val + op->nnz * vw);
            const int C = o.format == MK_FMT_CSR ? 64 : o.slice_height;
            mk::build_sell(op, rowptr, col, val, C, o.format == MK_FMT_CSR ? 1 : o.sigma);
        } catch (...) {
            delete op;
            throw;
        }
        *out = op;
    });
}


mk_status mk_autotune_spmv(mk_ctx* c, int n, const int64_t* rowptr, const int32_t* col, const double* val, int cplx,
                           const mk_solve_opts* base, mk_solve_opts* best, double* us) {
    if (!c || n <= 0 || !rowptr || !col || !val || !best) return MK_ERR_ARG;
    return guard(c, [&] {
        const mk_solve_opts b0 = resolve(base);
        MK_HIP(hipSetDevice(c->device));
        struct Cand { int format, C, sigma, block; };
        std::vector<Cand> cand = {{MK_FMT_CSR, 64, 1, 256}};
        for (int C : {64, 128, 256})
            for (int sigma : {1, 256})
                for (int block : {128, 256}) cand.push_back({MK_FMT_SELL, C, sigma, block});
        const size_t len = size_t(n) * (cplx ? 2 : 1);
        DevBuf<double> dx(len), dy(len);
        {
            std::vector<double> hx(len);
            for (size_t i = 0; i < len; i++) hx[i] = 1.0 + 1e-3 * double(i % 97);
            dx.upload(hx.data(), len, c->stream);
        }
        hipEvent_t e0, e1;
        MK_HIP(hipEventCreate(&e0));
        MK_HIP(hipEventCreate(&e1));
        double best_ms = HUGE_VAL;
        for (const Cand& k : cand) {
            mk_solve_opts o = b0;
            o.format = k.format;
            o.slice_height = k.C;
            o.sigma = k.sigma;
            o.block_size = k.block;
            o.precond = MK_PRECOND_NONE;
            mk_op* op = nullptr;
            const mk_status st = cplx ? mk_op_create_complex(c, n, rowptr, col, val, &o, &op)
                                      : mk_op_create_real(c, n, rowptr, col, val, &o, &op);
            if (st != MK_OK) continue;
            for (int w = 0; w < 3; w++) mk::spmv64(c, op, dx.p, dy.p, nullptr, mk::SPMV_PLAIN, nullptr, nullptr);
            constexpr int reps = 20;
            MK_HIP(hipEventRecord(e0, c->stream));
            for (int r = 0; r < reps; r++) mk::spmv64(c, op, dx.p, dy.p, nullptr, mk::SPMV_PLAIN, nullptr, nullptr);
            MK_HIP(hipEventRecord(e1, c->stream));
            MK_HIP(hipEventSynchronize(e1));
            float ms = 0.0f;
            MK_HIP(hipEventElapsedTime(&ms, e0, e1));
            mk_op_destroy(op);
            if (double(ms) / reps < best_ms) {
                best_ms = double(ms) / reps;
                *best = b0;
                best->format = k.format;
                best->slice_height = k.C;
                best->sigma = k.sigma;
                best->block_size = k.block;
            }
        }
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        MK_CHECK(best_ms < HUGE_VAL, MK_ERR_HIP, "autotune: no candidate layout could be built");
        if (us) *us = best_ms * 1e3;
    });
}

mk_status mk_op_create_real(mk_ctx* c, int n, const int64_t* rowptr, const int32_t* col, const double* val,
                            const mk_solve_opts* o, mk_op** op) {
    return op_create(c, n, rowptr, col, val, false, o, op);
}
mk_status mk_op_create_complex(mk_ctx* c, int n, const int64_t* rowptr, const int32_t* col, const double* val,
                               const mk_solve_opts* o, mk_op** op) {
    return op_create(c, n, rowptr, col, val, true, o, op);
}
mk_status mk_op_destroy(mk_op* op) {
    if (!op) return MK_ERR_ARG;
    (void)hipSetDevice(op->ctx->device);
    delete op;
    return MK_OK;
}

mk_status mk_op_apply(mk_ctx* c, const mk_op* A, const double* x, double* y) {
    if (!c || !A || !x || !y || A->ctx != c) return MK_ERR_ARG;
    return guard(c, [&] {
        MK_HIP(hipSetDevice(c->device));
        const size_t len = size_t(A->n) * (A->cplx ? 2 : 1);
        DevBuf<double> dx, dy(len);
        dx.upload(x, len, c->stream);
        mk::spmv64(c, A, dx.p, dy.p, nullptr, mk::SPMV_PLAIN, nullptr, nullptr);
        dy.download(y, len, c->stream);
        MK_HIP(hipStreamSynchronize(c->stream));
    });
}

mk_status mk_spd_solve(mk_ctx* c, const mk_op* A, const double* b, double* x, const mk_solve_opts* o_,
                       mk_stats* st) {
    if (!c || !A || !b || !x || A->ctx != c) return MK_ERR_ARG;
    mk_status inner = MK_OK;
    mk_status s = guard(c, [&] {
        const mk_solve_opts o = resolve(o_);
        MK_HIP(hipSetDevice(c->device));
        const double t0 = now_ms();
        mk::ensure_mg(c, A, o);
        const double t_setup = now_ms() - t0;
        mk::stats_reset(c);
        const size_t len = size_t(A->n) * (A->cplx ? 2 : 1);
        DevBuf<double> db, dx;
        db.upload(b, len, c->stream);
        dx.upload(x, len, c->stream);
        mk::SolveResult r = mk::solve_spd(c, A, db.p, dx.p, o);
        dx.download(x, len, c->stream);
        MK_HIP(hipStreamSynchronize(c->stream));
        if (st) {
            std::memset(st, 0, sizeof(*st));
            st->iterations = r.iters;
            st->refinements = r.refinements;
            st->solves = 1;
            st->rel_residual = r.rel_residual;
            mk::stats_fill(c, st);
            st->ms = now_ms() - t0;
            st->setup_ms = t_setup;
            st->mg_levels = A->mg ? mk::mg_levels(A->mg) : 0;
        }
        inner = r.status;
        if (inner != MK_OK) c->last_error = "mk_spd_solve: " + std::string(mk_status_str(inner));
    });
    return s != MK_OK ? s : inner;
}

__global__ void k_pack_cols(size_t n, int mw, int cnt, const double* src, double* dst) {
    const size_t e = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (e >= n * size_t(mw)) return;
    const size_t i = e / size_t(mw);
    const int j = int(e - i * size_t(mw));
    dst[e] = j < cnt ? src[size_t(j) * n + i] : 0.0;
}
__global__ void k_unpack_block(size_t n, int mw, int cnt, const double* src, double* dst) {
    const size_t e = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (e >= n * size_t(mw)) return;
    const size_t i = e / size_t(mw);
    const int j = int(e - i * size_t(mw));
    if (j < cnt) dst[size_t(j) * n + i] = src[e];
}

mk_status mk_spd_solve_many(mk_ctx* c, const mk_op* A, int m, const double* B, double* X, const mk_solve_opts* o_,
                            mk_stats* st) {
    if (!c || !A || m < 0 || (m > 0 && (!B || !X)) || A->ctx != c) return MK_ERR_ARG;
    mk_status inner = MK_OK;
    mk_status s = guard(c, [&] {
        const mk_solve_opts o = resolve(o_);
        MK_HIP(hipSetDevice(c->device));
        const double t0 = now_ms();
        mk::ensure_mg(c, A, o);
        const double t_setup = now_ms() - t0;
        mk::stats_reset(c);
        const size_t n = size_t(A->n), len = n * (A->cplx ? 2 : 1);
        int iters = 0, solves = 0;
        double worst = 0.0;
        // every precision mode batches the columns; MK_PREC_FP64 runs the V-cycle in FP64
        const bool batched = !A->cplx && A->mg && o.precond == MK_PRECOND_AMG;
        if (batched && m > 0) {
            const int mw = m > 32 ? 64 : m > 16 ? 32 : m > 8 ? 16 : 8;
            DevBuf<double> dB(n * size_t(std::min(m, mw))), dX(n * size_t(std::min(m, mw)));
            DevBuf<double> Bp(n * mw), Xp(n * mw);
            for (int b0 = 0; b0 < m && inner == MK_OK; b0 += mw) {
                const int cnt = std::min(mw, m - b0);
                dB.upload(B + size_t(b0) * n, size_t(cnt) * n, c->stream);
                hipLaunchKernelGGL(k_pack_cols, dim3(unsigned((n * mw + 255) / 256)), dim3(256), 0, c->stream, n, mw,
                                   cnt, (const double*)dB.p, Bp.p);
                mk::BlockResult r = mk::block_pcg(c, A, A->mg, mw, Bp.p, Xp.p, o.tol, o.maxiter, o.check_every,
                                                  o.precision == MK_PREC_FP64);
                hipLaunchKernelGGL(k_unpack_block, dim3(unsigned((n * mw + 255) / 256)), dim3(256), 0, c->stream, n,
                                   mw, cnt, (const double*)Xp.p, dX.p);
                dX.download(X + size_t(b0) * n, size_t(cnt) * n, c->stream);
                MK_HIP(hipStreamSynchronize(c->stream));
                iters += r.iters;
                solves += cnt;
                worst = std::max(worst, r.max_rel);
                inner = r.status;
            }
        } else {
            DevBuf<double> db, dx(len);
            for (int j = 0; j < m && inner == MK_OK; j++) {
                db.upload(B + size_t(j) * len, len, c->stream);
                MK_HIP(hipMemsetAsync(dx.p, 0, len * sizeof(double), c->stream));
                mk::SolveResult r = mk::solve_spd(c, A, db.p, dx.p, o);
                dx.download(X + size_t(j) * len, len, c->stream);
                MK_HIP(hipStreamSynchronize(c->stream));
                iters += r.iters;
                solves++;
                worst = std::max(worst, r.rel_residual);
                inner = r.status;
            }
        }
        if (st) {
            std::memset(st, 0, sizeof(*st));
            st->iterations = iters;
            st->solves = solves;
            st->rel_residual = worst;
            mk::stats_fill(c, st);
            st->ms = now_ms() - t0;
            st->setup_ms = t_setup;
            st->mg_levels = A->mg ? mk::mg_levels(A->mg) : 0;
        }
        if (inner != MK_OK) c->last_error = "mk_spd_solve_many: " + std::string(mk_status_str(inner));
    });
    return s != MK_OK ? s : inner;
}

mk_status mk_field_laplacian(mk_ctx* c, int nF, int nH, const int32_t* face_halfs, const int32_t* twin,
                             const int32_t* half_face, const double* farg, int N, const mk_solve_opts* o_,
                             mk_op** out) {
    if (!c || !out || nF <= 0 || nH <= 0 || !face_halfs || !twin || !half_face || !farg || N <= 0) return MK_ERR_ARG;
    return guard(c, [&] {
        const mk_solve_opts o = resolve(o_);
        MK_HIP(hipSetDevice(c->device));
        for (int64_t k = 0; k < int64_t(nF) * 3; k++)
            MK_CHECK(face_halfs[k] >= 0 && face_halfs[k] < nH, MK_ERR_ARG, "face_halfs out of range");
        for (int h = 0; h < nH; h++) {
            MK_CHECK(twin[h] >= 0 && twin[h] < nH, MK_ERR_ARG, "twin out of range");
            MK_CHECK(half_face[h] >= -1 && half_face[h] < nF, MK_ERR_ARG, "half_face out of range");
        }
        const int C = o.slice_height;
        MK_CHECK(C % 64 == 0 && C > 0, MK_ERR_ARG, "slice_height");
        auto* op = new mk_op();
        try {
            op->ctx = c;
            op->n = nF;
            op->m = nF;
            op->label = "field_L";
            op->n_short = nF;
            op->nlong = 0;
            op->cplx = true;
            op->C = C;
            op->sigma = 1;
            op->block = o.block_size;
            MK_CHECK(op->block % 64 == 0 && op->block >= 64 && op->block <= 256, MK_ERR_ARG,
                     "block_size must be 64, 128 or 256");
            const int ns = (nF + C - 1) / C;
            op->nslices = ns;
            op->stored = int64_t(ns) * 4 * C;
            DevBuf<int> dfh, dtw, dhf;
            DevBuf<double> dfa;
            dfh.upload(face_halfs, size_t(nF) * 3, c->stream);
            dtw.upload(twin, size_t(nH), c->stream);
            dhf.upload(half_face, size_t(nH), c->stream);
            dfa.upload(farg, size_t(nH), c->stream);
            op->col.alloc(size_t(op->stored));
            op->val64.alloc(size_t(op->stored) * 2);
            op->val32.alloc(size_t(op->stored) * 2);
            op->slice_ptr.alloc(size_t(ns) + 1);
            op->slice_w.alloc(size_t(ns));
            op->dinv64.alloc(size_t(nF));
            op->dinv32.alloc(size_t(nF));
            DevBuf<int> cnt(static_cast<size_t>(nF));
            // padding rows of the last slice: zero values on a valid column
            hipLaunchKernelGGL(k_fill_i32, dim3(nblk(op->stored)), dim3(256), 0, c->stream, size_t(op->stored), 0,
                               op->col.p);
            MK_HIP(hipMemsetAsync(op->val64.p, 0, size_t(op->stored) * 2 * sizeof(double), c->stream));
            hipLaunchKernelGGL(k_slice_ptr, dim3(nblk(size_t(ns) + 1)), dim3(256), 0, c->stream, ns, int64_t(4),
                               int64_t(C), op->slice_ptr.p, op->slice_w.p);
            hipLaunchKernelGGL(k_field_laplacian, dim3(nblk(nF)), dim3(256), 0, c->stream, nF, C, dfh.p, dtw.p, dhf.p,
                               dfa.p, N, op->col.p, op->val64.p, cnt.p);
            hipLaunchKernelGGL(k_to_f32, dim3(nblk(size_t(op->stored) * 2)), dim3(256), 0, c->stream,
                               size_t(op->stored) * 2, op->val64.p, op->val32.p);
            hipLaunchKernelGGL(k_dinv_from_diag, dim3(nblk(nF)), dim3(256), 0, c->stream, nF, C, op->col.p,
                               op->val64.p, op->dinv64.p, op->dinv32.p);
            MK_HIP(hipGetLastError());
            // the same rows on the host for the multigrid set-up, while the device assembles
            // (spares the set-up an export of the device operator)
            if (o.precond == MK_PRECOND_AMG) {
                std::vector<int64_t>& rp = op->mg_ptr;
                rp.assign(size_t(nF) + 1, 0);
#pragma omp parallel for schedule(static, 4096)
                for (int f = 0; f < nF; f++) {
                    int m = 1;
                    for (int k = 0; k < 3; k++) m += half_face[twin[face_halfs[3 * f + k]]] >= 0 ? 1 : 0;
                    rp[size_t(f) + 1] = m;
                }
                for (int f = 0; f < nF; f++) rp[size_t(f) + 1] += rp[size_t(f)];
                op->mg_col.resize(size_t(rp[size_t(nF)]));
                op->mg_val.resize(size_t(rp[size_t(nF)]) * 2);
                const double sgn = (N % 2 == 0) ? -1.0 : 1.0;
#pragma omp parallel for schedule(static, 4096)
                for (int f = 0; f < nF; f++) {
                    int cc[4];
                    double vr[4], vi[4];
                    int m = 0;
                    double diag = 0.0;
                    for (int k = 0; k < 3; k++) {
                        const int h = face_halfs[3 * f + k], tw = twin[h], g = half_face[tw];
                        if (g < 0) continue;
                        diag += 1.0;
                        const double a = double(N) * (farg[h] - farg[tw]);
                        cc[m] = g;
                        vr[m] = sgn * std::cos(a);
                        vi[m] = sgn * std::sin(a);
                        m++;
                    }
                    cc[m] = f;
                    vr[m] = diag + 1e-9;
                    vi[m] = 0.0;
                    m++;
                    for (int i = 1; i < m; i++)
                        for (int j = i; j > 0 && cc[j - 1] > cc[j]; j--) {
                            std::swap(cc[j - 1], cc[j]);
                            std::swap(vr[j - 1], vr[j]);
                            std::swap(vi[j - 1], vi[j]);
                        }
                    const int64_t b = rp[size_t(f)];
                    for (int j = 0; j < m; j++) {
                        op->mg_col[size_t(b + j)] = cc[j];
                        op->mg_val[2 * size_t(b + j)] = vr[j];
                        op->mg_val[2 * size_t(b + j) + 1] = vi[j];
                    }
                }
            }
            std::vector<int> hc(static_cast<size_t>(nF));
            cnt.download(hc.data(), hc.size(), c->stream);
            MK_HIP(hipStreamSynchronize(c->stream));
            op->nnz = 0;
            for (int v : hc) op->nnz += v;
            const double vw = 2.0;
            op->bytes64 = double(op->nnz) * (8.0 * vw + 4.0) + double(ns) * 12.0 + double(nF) * 8.0 * vw * 2.0;
            op->bytes32 = double(op->nnz) * (4.0 * vw + 4.0) + double(ns) * 12.0 + double(nF) * 4.0 * vw * 2.0;
        } catch (...) {
            delete op;
            throw;
        }
        *out = op;
    });
}

mk_status mk_op_export_csr(mk_ctx* c, const mk_op* A, int64_t* nnz, int64_t* rowptr, int32_t* col, double* val) {
    if (!c || !A || !nnz || A->ctx != c) return MK_ERR_ARG;
    return guard(c, [&] {
        std::vector<int64_t> rp;
        std::vector<int> cc;
        std::vector<double> vv;
        mk::host_csr(c, A, rp, cc, vv);
        *nnz = rp[A->n];
        if (rowptr) std::copy(rp.begin(), rp.end(), rowptr);
        if (col) std::copy(cc.begin(), cc.end(), col);
        if (val) std::copy(vv.begin(), vv.end(), val);
    });
}

mk_status mk_field_smallest_eig(mk_ctx* c, const mk_op* L, const double* mass, const double* u0, int n_iter,
                                const mk_solve_opts* o_, double* x_out, mk_stats* st) {
    if (!c || !L || !mass || !u0 || !x_out || n_iter < 0 || !L->cplx || L->ctx != c) return MK_ERR_ARG;
    mk_status inner = MK_OK;
    mk_status s = guard(c, [&] {
        const mk_solve_opts o = resolve(o_);
        MK_HIP(hipSetDevice(c->device));
        double t0 = now_ms();
        mk::ensure_mg(c, L, o);
        const double t_setup = now_ms() - t0;
        t0 = now_ms();
        mk::stats_reset(c);
        const int n = L->n;
        const size_t len = size_t(n) * 2;
        DevBuf<double> dm, du, dx, db, dW, dLW, dpart, dy;
        dm.upload(mass, size_t(n), c->stream);
        du.upload(u0, len, c->stream);
        dx.alloc(len);
        db.alloc(len);
        dW.alloc(len * kWin);
        dLW.alloc(len * kWin);
        const unsigned gw = mk::rgrid(size_t(n));
        dpart.alloc(size_t(gw) * kWinDots + kWinDots);
        dy.alloc(2 * kWin);
        double* sc = c->scal.p;
        int total_it = 0, total_ref = 0;
        double last_rel = 0.0;
        // Galerkin matrix G = W^H L W of the window (slot order), kept across iterations
        std::vector<std::complex<double>> G(static_cast<size_t>(kWin) * kWin), g(static_cast<size_t>(kWin));
        double hd[kWinDots], hy[2 * kWin];
        for (int it = 0; it < n_iter; it++) {
            // b = M u
            mk::mass_mul(c, n, dm.p, du.p, db.p);
            // The iterate enters the window (slot s) with L u, then the start is the Galerkin
            // solution on span(W): x0 = W (W^H L W)^+ W^H b.  With the window {u} alone this is
            // the Rayleigh-quotient start u (u^H M u) / (u^H L u); the last kWin iterates
            // span the slowly converging eigen-directions, so x0 starts PCG many digits closer.
            const int s = it % kWin, nw = std::min(it + 1, kWin);
            double* ws = dW.p + len * size_t(s);
            double* lws = dLW.p + len * size_t(s);
            MK_HIP(hipMemcpyAsync(ws, du.p, len * sizeof(double), hipMemcpyDeviceToDevice, c->stream));
            mk::spmv64(c, L, du.p, lws, nullptr, mk::SPMV_PLAIN, nullptr, nullptr);
            hipLaunchKernelGGL(k_window_dots, dim3(gw), dim3(256), 0, c->stream, n, nw, (const double*)dW.p,
                               (const double*)lws, (const double*)db.p, dpart.p);
            hipLaunchKernelGGL(k_window_sum, dim3(1), dim3(64 * kWinDots), 0, c->stream, int(gw), (const double*)dpart.p,
                               dpart.p + size_t(gw) * kWinDots);
            MK_HIP(hipMemcpyAsync(hd, dpart.p + size_t(gw) * kWinDots, sizeof(hd), hipMemcpyDeviceToHost,
                                  c->stream));
            MK_HIP(hipStreamSynchronize(c->stream));
            for (int j = 0; j < nw; j++) {
                const std::complex<double> gjs(hd[4 * j], hd[4 * j + 1]);  // w_j^H L w_s
                G[size_t(j) * kWin + s] = gjs;
                G[size_t(s) * kWin + j] = std::conj(gjs);
                g[size_t(j)] = {hd[4 * j + 2], hd[4 * j + 3]};            // w_j^H b
            }
            G[size_t(s) * kWin + s] = hd[4 * s];
            MK_CHECK(hd[4 * s] == hd[4 * s] && hd[4 * s] > 0.0, MK_ERR_NAN, "u^H L u not positive");
            std::vector<std::complex<double>> Gn(size_t(nw) * nw), gn(g.begin(), g.begin() + nw);
            for (int i = 0; i < nw; i++)
                for (int j = 0; j < nw; j++) Gn[size_t(i) * nw + j] = G[size_t(i) * kWin + j];
            const std::vector<std::complex<double>> y = mk::herm_pinv_solve(nw, Gn, gn);
            for (int j = 0; j < nw; j++) { hy[2 * j] = y[size_t(j)].real(); hy[2 * j + 1] = y[size_t(j)].imag(); }
            MK_HIP(hipMemcpyAsync(dy.p, hy, 2 * size_t(nw) * sizeof(double), hipMemcpyHostToDevice, c->stream));
            hipLaunchKernelGGL(k_window_combine, dim3(nblk(size_t(n))), dim3(256), 0, c->stream, n, nw,
                               (const double*)dW.p, (const double*)dy.p, dx.p);
            // Each solve aims at a tenth of the required relative residual (as far as the FP64
            // floor of the true residual allows): the iteration keeps every solve's error in the
            // field, and the parameterization downstream amplifies field error about a
            // thousandfold into uv.
            mk::SolveResult r = mk::solve_spd(c, L, db.p, dx.p, o, nullptr, 0.1 * o.tol);
            total_it += r.iters;
            total_ref += r.refinements;
            last_rel = r.rel_residual;
            if (it % 10 == 9) MK_LOG("[field] inverse iteration %d: %d PCG iterations so far\n", it + 1, total_it);
            if (r.status != MK_OK) {
                inner = r.status;
                c->last_error = "inverse iteration " + std::to_string(it) + ": " + mk_status_str(r.status);
                break;
            }
            // s = sqrt(|x^H M x|); x /= s; u = x
            mk::mass_dot(c, n, dm.p, dx.p, sc + 12);
            mk::scale_dev64(c, len, sc + 12, 1, dx.p);
            MK_HIP(hipMemcpyAsync(du.p, dx.p, len * sizeof(double), hipMemcpyDeviceToDevice, c->stream));
        }
        dx.download(x_out, len, c->stream);
        MK_HIP(hipStreamSynchronize(c->stream));
        if (st) {
            std::memset(st, 0, sizeof(*st));
            st->iterations = total_it;
            st->refinements = total_ref;
            st->solves = n_iter;
            st->rel_residual = last_rel;
            mk::stats_fill(c, st);
            st->ms = now_ms() - t0;
            st->setup_ms = t_setup;
            st->mg_levels = L->mg ? mk::mg_levels(L->mg) : 0;
        }
    });
    return s != MK_OK ? s : inner;
}

}  // extern "C"
