// mk_iterative_rounding: the IGM solve of metriko (iter_rounding.h:33-113) on the GPU.
//
// Host: constraint kernel basis (FullPivLU restatement), UFull, X2F = G2*UFull and the
// Poisson matrix E = length*X2F^T X2F (sparse products), control flow of the LM passes.
// Device: every linear solve (PCG, FP64 or FP32 + FP64 refinement), the energies the
// LM acceptance tests read, the fixed-row dot products and the FP64 rounding values.
//
// Each Levenberg-Marquardt pass of the reference runs with lambda = 0 on an affine
// residual, so its direction is d = x* - x_prev with x* the minimiser of the quadratic
// objective.  x* is computed in an exactly equivalent, well-conditioned form:
//  * naive pass: f eliminated (diagonal Schur complement); x* = the KKT solution;
//  * rounding passes: A = A0 + wK^2 U^T U, A0 = X2F^T X2F + wC^2 I fixed for the loop,
//    x* = xt + Z (wK^-2 I + U Z)^{-1} (r - U xt), xt = A0^{-1} b0, Z = A0^{-1} U^T
//    (Sherman-Morrison-Woodbury; no 1e10-weighted quantity is ever formed).
#include <algorithm>
#include <map>
#include <memory>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstring>
#include <numeric>
#include "mk_internal.h"
#include "igm_state.h"
#include "metriko/core/common/sparse.h"
#include "metriko/core/solver/matrix_ops.h"

using mk::DevBuf;
using mk::Error;
using metriko::Csr;
using metriko::SprsD;
using metriko::TripD;

namespace {

double now_ms() {
    return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now().time_since_epoch()).count();
}
inline unsigned nblk(size_t n, int bs = 256) { return unsigned((n + bs - 1) / bs); }

// ---- device helpers ----------------------------------------------------------------------
__device__ __forceinline__ double wsum(double v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}
// one block per sparse row: out[i] = row_i . x   (deterministic block reduction)
__global__ void k_rowset_dot(int nrows, const int64_t* ptr, const int* col, const double* val, const double* x,
                             double* out) {
    const int i = blockIdx.x;
    if (i >= nrows) return;
    double s = 0.0;
    for (int64_t k = ptr[i] + threadIdx.x; k < ptr[i + 1]; k += blockDim.x) s += val[k] * x[col[k]];
    __shared__ double sh[4];
    s = wsum(s);
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = 0.0;
        for (int w = 0; w < int(blockDim.x >> 6); w++) t += sh[w];
        out[i] = t;
    }
}
// y = 0; y[col] = val for row i
__global__ void k_scatter_row(int64_t b, int64_t e, const int* col, const double* val, double* y) {
    const int64_t k = b + int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (k < e) y[col[k]] = val[k];
}
__global__ void k_zero(size_t n, double* y) {
    const size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) y[i] = 0.0;
}
// y = x + sum_j c[j] * Z[:, idx[j]]   (Z column-major, k rows)
__global__ void k_combine(size_t k, int r, const double* x, const double* Z, const int* idx, const double* c,
                          double* y) {
    const size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= k) return;
    double s = x[i];
    for (int j = 0; j < r; j++) s += c[j] * Z[size_t(idx[j]) * k + i];
    y[i] = s;
}
// packed right-hand sides: B[col * mw + j] = row sel[j] of a RowSet (sel[j] < 0: empty column)
__global__ void k_scatter_rows_packed(int mw, const int* sel, const int64_t* ptr, const int* col, const double* val,
                                      double* B) {
    const int j = blockIdx.x;
    const int p = sel[j];
    if (p < 0) return;
    for (int64_t e = ptr[p] + threadIdx.x; e < ptr[p + 1]; e += blockDim.x) B[size_t(col[e]) * mw + j] = val[e];
}
// Z[:, sel[j]] = X[:, j] for the packed n x mw block X
__global__ void k_unpack_cols(size_t n, int mw, const int* sel, const double* X, double* Z) {
    const size_t e = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (e >= n * size_t(mw)) return;
    const size_t i = e / size_t(mw);
    const int j = int(e - i * size_t(mw));
    const int p = sel[j];
    if (p >= 0) Z[size_t(p) * n + i] = X[e];
}
// UZ[c * nc + p] = row c of the RowSet . Z[:, p]
__global__ void k_uz(int nc, size_t n, const int64_t* ptr, const int* col, const double* val, const double* Z,
                     double* UZ) {
    const size_t t = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (t >= size_t(nc) * size_t(nc)) return;
    const size_t c = t / size_t(nc), p = t - c * size_t(nc);
    double s = 0.0;
    for (int64_t e = ptr[c]; e < ptr[c + 1]; e++) s += val[e] * Z[p * n + size_t(col[e])];
    UZ[t] = s;
}
// y = a*x + b*z
__global__ void k_lin2(size_t n, double a, const double* x, double b, const double* z, double* y) {
    const size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) y[i] = a * x[i] + b * z[i];
}
// partial sums of w * (s*x - t)^2 over n entries (t may be null -> 0); deterministic two-pass
__global__ void k_sqdiff_partial(size_t n, double s, const double* x, const double* t, double* part) {
    double acc = 0.0;
    for (size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += size_t(gridDim.x) * blockDim.x) {
        const double d = s * x[i] - (t ? t[i] : 0.0);
        acc += d * d;
    }
    __shared__ double sh[4];
    acc = wsum(acc);
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = acc;
    __syncthreads();
    if (threadIdx.x == 0) {
        double q = 0.0;
        for (int w = 0; w < int(blockDim.x >> 6); w++) q += sh[w];
        part[blockIdx.x] = q;
    }
}
__global__ void k_sum_partials(int np, const double* part, double* out) {
    double acc = 0.0;
    for (int i = threadIdx.x; i < np; i += blockDim.x) acc += part[i];
    __shared__ double sh[4];
    acc = wsum(acc);
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = acc;
    __syncthreads();
    if (threadIdx.x == 0) {
        double q = 0.0;
        for (int w = 0; w < int(blockDim.x >> 6); w++) q += sh[w];
        out[0] = q;
    }
}

// ---- device sparse objects -----------------------------------------------------------------
struct RowSet {  // a small set of sparse rows (CSR) kept on the device
    int nrows = 0;
    std::vector<int64_t> hptr{0};
    std::vector<int> hcol;
    std::vector<double> hval;
    DevBuf<int64_t> ptr;
    DevBuf<int> col;
    DevBuf<double> val;
    const std::vector<int>* cmap = nullptr;  // device numbering of the columns (null: the same)
    void push(const SprsD& M, int row) {
        for (int64_t k = M.ptr[row]; k < M.ptr[row + 1]; k++) {
            hcol.push_back(cmap && !cmap->empty() ? (*cmap)[size_t(M.col[k])] : M.col[k]);
            hval.push_back(M.val[k]);
        }
        hptr.push_back(int64_t(hcol.size()));
        nrows++;
    }
    void sync(hipStream_t s) {
        ptr.upload(hptr.data(), hptr.size(), s);
        col.upload(hcol.data(), std::max<size_t>(hcol.size(), 1), s);
        val.upload(hval.data(), std::max<size_t>(hval.size(), 1), s);
    }
};

struct Ctx {
    mk_ctx* c;
    mk_solve_opts o;
    mk_stats acc{};
    DevBuf<double> part;
    explicit Ctx(mk_ctx* c_, const mk_solve_opts& o_) : c(c_), o(o_), part(4096) {}
    hipStream_t st() const { return c->stream; }
    void sync() { MK_HIP(hipStreamSynchronize(c->stream)); }
    // ||s*x - t||^2 on the device -> host
    double sqdiff(size_t n, double s, const double* x, const double* t) {
        const unsigned g = std::min<unsigned>(std::max(1u, nblk(n)), 1024u);
        hipLaunchKernelGGL(k_sqdiff_partial, dim3(g), dim3(256), 0, st(), n, s, x, t, part.p);
        hipLaunchKernelGGL(k_sum_partials, dim3(1), dim3(256), 0, st(), int(g), part.p, part.p + 2048);
        double h = 0.0;
        MK_HIP(hipMemcpyAsync(&h, part.p + 2048, sizeof(double), hipMemcpyDeviceToHost, st()));
        sync();
        return h;
    }
    void rowdots(const RowSet& R, const double* x, double* out_dev) {
        if (R.nrows == 0) return;
        hipLaunchKernelGGL(k_rowset_dot, dim3(R.nrows), dim3(256), 0, st(), R.nrows, R.ptr.p, R.col.p, R.val.p, x,
                           out_dev);
    }
    mk::SolveResult solve(const mk_op* A, const double* b, double* x, mk::Mg* pre = nullptr) {
        mk::SolveResult r = mk::solve_spd(c, A, b, x, o, pre);
        acc.iterations += r.iters;
        acc.refinements += r.refinements;
        acc.solves += 1;
        acc.rel_residual = r.rel_residual;
        if (r.status != MK_OK) throw Error(r.status, "IGM solve failed: " + std::string(mk_status_str(r.status)));
        return r;
    }
};

// Values of s A + d I for a matrix whose rows all hold their diagonal entry (E = X2F^T X2F
// of a variable set every column of X2F touches); OpenMP over rows.
metriko::HVec<double> scaled_shifted(const SprsD& A, double s, double d) {
    const int n = A.rows();
    metriko::HVec<double> v(A.val.size());
    std::atomic<int> missing{0};
#pragma omp parallel for schedule(static, 4096)
    for (int i = 0; i < n; i++) {
        bool diag = false;
        for (int64_t q = A.ptr[i]; q < A.ptr[i + 1]; q++) {
            v[size_t(q)] = s * A.val[size_t(q)];
            if (A.col[size_t(q)] == i) { v[size_t(q)] += d; diag = true; }
        }
        if (!diag) missing.store(1, std::memory_order_relaxed);
    }
    if (missing.load()) throw Error(MK_ERR_SINGULAR, "rounding operator: a variable no gradient row touches");
    return v;
}

// rectangular operator (no multigrid, no diagonal): hybrid SELL upload of a host CSR
// device operator (no host copy kept) from CSR arrays, values times scale
mk_op* make_op_arrays(mk_ctx* c, int rows, int cols, const int64_t* ptr, const int* col, const double* val,
                      const mk_solve_opts& o, double scale = 1.0, const std::vector<int>* rsrc = nullptr,
                      const std::vector<int>* cmap = nullptr) {
    auto* op = new mk_op();
    try {
        op->ctx = c;
        op->n = rows;
        op->m = cols;
        op->cplx = false;
        op->nnz = ptr[rows];
        op->block = o.block_size;
        const bool csr = o.format == MK_FMT_CSR;
        mk::build_sell(op, ptr, col, val, csr ? 64 : o.slice_height, csr ? 1 : o.sigma, scale,
                       rsrc && !rsrc->empty() ? rsrc->data() : nullptr, cmap && !cmap->empty() ? cmap->data() : nullptr);
    } catch (...) {
        delete op;
        throw;
    }
    return op;
}
mk_op* make_rect_op(mk_ctx* c, const SprsD& A, const mk_solve_opts& o, double scale = 1.0,
                    const std::vector<int>* rsrc = nullptr, const std::vector<int>* cmap = nullptr) {
    return make_op_arrays(c, A.rows(), A.cols(), A.ptr.data(), A.col.data(), A.val.data(), o, scale, rsrc, cmap);
}

// The A0 system's unknowns are kept on the device in the locality order of A0's multigrid
// aggregates (the host keeps the caller's numbering): in the caller's vertex order neighbouring
// unknowns of a subdivided mesh lie hundreds of thousands of indices apart, and every product
// on A0 then gathers from all over its vector.  The operators' rows keep their entries' order,
// so every product computes the same numbers.  METRIKO_IGM_ORDER=0 turns it off, =2 on for any
// size; by default it is on from 65,536 unknowns.  The local-injectivity passes keep the
// caller's order.
bool igm_order_wanted(int k, const mk_igm_flags* flags) {
    if (flags->local_injectivity) return false;
    const char* e = std::getenv("METRIKO_IGM_ORDER");
    const int m = e ? std::atoi(e) : 1;
    return m == 2 || (m == 1 && k >= 65536);
}
// out[pos[i]] = h[i] (device order) / out[i] = d[pos[i]] (the caller's order); pos empty: copies
std::vector<double> to_dev_order(const std::vector<int>& pos, const double* h, size_t n) {
    std::vector<double> out(n);
    if (pos.empty()) { std::copy(h, h + n, out.begin()); return out; }
    for (size_t i = 0; i < n; i++) out[size_t(pos[i])] = h[i];
    return out;
}
void from_dev_order(const std::vector<int>& pos, const std::vector<double>& d, double* out) {
    if (pos.empty()) { std::copy(d.begin(), d.end(), out); return; }
    for (size_t i = 0; i < d.size(); i++) out[i] = d[size_t(pos[i])];
}
struct OpHolder {
    mk_op* op = nullptr;
    ~OpHolder() { if (op) mk_op_destroy(op); }
};

// dense solve with partial pivoting (small r x r)
std::vector<double> dense_solve(std::vector<double> M, std::vector<double> b, int r) {
    std::vector<int> piv(static_cast<size_t>(r));
    for (int k = 0; k < r; k++) {
        int p = k;
        for (int i = k + 1; i < r; i++) if (std::abs(M[size_t(i * r + k)]) > std::abs(M[size_t(p * r + k)])) p = i;
        if (M[size_t(p * r + k)] == 0.0) throw Error(MK_ERR_SINGULAR, "rounding: singular penalty system");
        if (p != k) { for (int j = 0; j < r; j++) std::swap(M[size_t(k * r + j)], M[size_t(p * r + j)]); std::swap(b[k], b[p]); }
        for (int i = k + 1; i < r; i++) {
            const double l = M[size_t(i * r + k)] / M[size_t(k * r + k)];
            if (l == 0.0) continue;
            for (int j = k; j < r; j++) M[size_t(i * r + j)] -= l * M[size_t(k * r + j)];
            b[i] -= l * b[k];
        }
    }
    std::vector<double> x(static_cast<size_t>(r));
    for (int i = r - 1; i >= 0; i--) {
        double s = b[i];
        for (int j = i + 1; j < r; j++) s -= M[size_t(i * r + j)] * x[j];
        x[i] = s / M[size_t(i * r + i)];
    }
    return x;
}

// LMSolver::solve (levenberg_marquardt.h:140-231) with lambda = 0, verbose = false.
// dir(prevx) writes the direction into d (device); energy(x) returns ||E(x)||^2.
// Returns 1/0 like the reference; x_out receives the solution (device, may alias prevx buffer).
// post(x) -> true ends the solve after an iteration (ST->post_iteration, levenberg_marquardt.h:221-224).
template <typename Dir, typename Energy, typename Post>
bool lm_solve(Ctx& K, size_t n, double* x0, double* x_out, int maxIter, double funcTol, Dir&& dir, Energy&& energy,
              Post&& post, DevBuf<double>& prevx, DevBuf<double>& d, DevBuf<double>& trial) {
    prevx.ensure(n); d.ensure(n); trial.ensure(n);
    MK_HIP(hipMemcpyAsync(prevx.p, x0, n * sizeof(double), hipMemcpyDeviceToDevice, K.st()));
    int it = 0;
    bool ret = false;
    for (;;) {
        dir(prevx.p, d.p);
        const double dn = std::sqrt(K.sqdiff(n, 1.0, d.p, nullptr));
        if (dn < funcTol) { ret = true; break; }
        const double e_prev = energy(prevx.p);
        hipLaunchKernelGGL(k_lin2, dim3(nblk(n)), dim3(256), 0, K.st(), n, 1.0, prevx.p, 1.0, d.p, trial.p);
        const double e_new = energy(trial.p);
        if (e_prev > e_new) {
            MK_HIP(hipMemcpyAsync(prevx.p, trial.p, n * sizeof(double), hipMemcpyDeviceToDevice, K.st()));
            if (std::abs(e_prev - e_new) < funcTol) { ret = false; break; }
        }
        if (post(prevx.p)) { ret = true; break; }
        it++;
        if (!(it <= maxIter)) { ret = false; break; }
    }
    MK_HIP(hipMemcpyAsync(x_out, prevx.p, n * sizeof(double), hipMemcpyDeviceToDevice, K.st()));
    return ret;
}

// Near-null vectors of E = X2F^T X2F for the multigrid setup: the constant parameterisations
// (u, v) = (1, 0) and (0, 1).  RosyParameterization lays the variables out as n dofs per
// vertex (component = index mod n) followed by n dofs per seam transition, and the transition
// dofs are exactly integerIdcs (setup.ipp:163-170).  The vertex dofs take the constant; the
// transition dofs solve min ||G2 y||^2 + ||Cfull y||^2 (consistent: the constant map is
// seamless; transitions absent from G2 are fixed by the constraints), normal equations by
// Jacobi-CG on the host.  The
// compressed coordinates are read off UFull: identity rows for nonPart, and the unit rows
// FullPivLU::kernel puts at each kernel column's free variable.  Returns n*k values
// (column-major), or nothing when a step does not apply or the vectors are not null to
// round-off; only the multigrid convergence rate depends on them.
// Non-owning CSR view (the caller's G2 arrays are used in place).
struct CsrView {
    int nr = 0, nc = 0;
    const int64_t* ptr = nullptr;
    const int* col = nullptr;
    const double* val = nullptr;
    int rows() const { return nr; }
    int cols() const { return nc; }
    static CsrView of(const SprsD& A) { return {A.rows(), A.cols(), A.ptr.data(), A.col.data(), A.val.data()}; }
};

// X2F = G2 * UFull from the caller's G2 arrays.  UFull rows hold one entry (unconstrained
// DOFs) or a USmall row (the few constrained DOFs), so an X2F row is the merge of a handful of
// products: expanded, sorted by column (stable), duplicates summed in order and exact zeros
// dropped (the SpGEMM + pruned() result).  Two OpenMP passes (count, fill); a G2 column out
// of range is reported as MK_ERR_ARG.
SprsD x2f_product(const CsrView& G2, const SprsD& UF, int npn) {
    const int rows = G2.nr;
    SprsD X(rows, UF.cols());
    X.ptr.assign(size_t(rows) + 1, 0);
    std::atomic<int> bad{0};
    // the unconstrained DOFs' UFull rows are single entries onto distinct, ascending columns
    // [0, npn): a G2 row that touches only those is relabelled in place (its order is kept, its
    // products are the ones the merge would form, exact zeros dropped as the merge does)
    std::vector<int> one(static_cast<size_t>(G2.nc));
#pragma omp parallel for schedule(static, 65536) if (G2.nc > 100000)
    for (int c = 0; c < G2.nc; c++) {
        const int64_t b = UF.ptr[c];
        one[size_t(c)] = (UF.ptr[c + 1] - b == 1 && UF.col[b] < npn) ? UF.col[b] : -1;
    }
    auto simple = [&](int r) {  // only such columns, in ascending order
        int prev = -1;
        for (int64_t q = G2.ptr[r]; q < G2.ptr[r + 1]; q++) {
            const int c = G2.col[q];
            if (c < 0 || c >= G2.nc || one[size_t(c)] <= prev) return false;
            prev = one[size_t(c)];
        }
        return true;
    };
    auto expand = [&](int r, std::vector<std::pair<int, double>>& buf) {
        buf.clear();
        for (int64_t q = G2.ptr[r]; q < G2.ptr[r + 1]; q++) {
            const int c = G2.col[q];
            if (c < 0 || c >= G2.nc) { bad.store(1, std::memory_order_relaxed); continue; }
            for (int64_t t = UF.ptr[c]; t < UF.ptr[c + 1]; t++) buf.emplace_back(UF.col[t], G2.val[q] * UF.val[t]);
        }
        for (size_t i = 1; i < buf.size(); i++) {  // stable insertion sort (a few entries)
            const auto e = buf[i];
            size_t j = i;
            while (j > 0 && buf[j - 1].first > e.first) { buf[j] = buf[j - 1]; j--; }
            buf[j] = e;
        }
        size_t w = 0;
        for (size_t i = 0; i < buf.size();) {
            const int cc = buf[i].first;
            double sum = 0.0;
            while (i < buf.size() && buf[i].first == cc) sum += buf[i++].second;
            if (sum != 0.0) buf[w++] = {cc, sum};
        }
        buf.resize(w);
    };
#pragma omp parallel
    {
        std::vector<std::pair<int, double>> buf;
#pragma omp for schedule(static, 4096)
        for (int r = 0; r < rows; r++) {
            if (simple(r)) {
                int64_t m = 0;
                for (int64_t q = G2.ptr[r]; q < G2.ptr[r + 1]; q++)
                    m += (G2.val[q] * UF.val[UF.ptr[G2.col[q]]] != 0.0) ? 1 : 0;
                X.ptr[size_t(r) + 1] = m;
                continue;
            }
            expand(r, buf);
            X.ptr[size_t(r) + 1] = int64_t(buf.size());
        }
    }
    if (bad.load()) throw Error(MK_ERR_ARG, "G2 column out of range");
    for (int r = 0; r < rows; r++) X.ptr[size_t(r) + 1] += X.ptr[size_t(r)];
    X.col.resize(size_t(X.ptr[size_t(rows)]));
    X.val.resize(size_t(X.ptr[size_t(rows)]));
#pragma omp parallel
    {
        std::vector<std::pair<int, double>> buf;
#pragma omp for schedule(static, 4096)
        for (int r = 0; r < rows; r++) {
            int64_t w = X.ptr[size_t(r)];
            if (simple(r)) {
                for (int64_t q = G2.ptr[r]; q < G2.ptr[r + 1]; q++) {
                    const int c = G2.col[q];
                    const double v = G2.val[q] * UF.val[UF.ptr[c]];
                    if (v != 0.0) { X.col[size_t(w)] = one[size_t(c)]; X.val[size_t(w)] = v; w++; }
                }
                continue;
            }
            expand(r, buf);
            for (const auto& e : buf) { X.col[size_t(w)] = e.first; X.val[size_t(w)] = e.second; w++; }
        }
    }
    return X;
}

std::vector<double> near_null(const CsrView& G2, const CsrView& Cf, const SprsD& X2F, int n, const int* intIdx,
                              int nint, const std::vector<int>& nonPart, const std::vector<int>& uniqueJ,
                              const metriko::MatXd& US) {
    const int nfull = G2.cols(), k = X2F.cols(), npn = int(nonPart.size());
    if (n < 1 || n > 4) return {};
    const double t_nn = now_ms();
    std::vector<int> tpos(size_t(nfull), -1);
    for (int i = 0; i < nint; i++) if (tpos[intIdx[i]] < 0) tpos[intIdx[i]] = i;
    // free-variable (unit) row of each kernel column
    const int dimker = US.rows() ? US.cols() : 0;
    std::vector<int> freeRow(size_t(dimker), -1);
    for (int r = 0; r < US.rows(); r++) {
        int nzj = -1, cnt = 0;
        for (int j = 0; j < dimker && cnt < 2; j++) if (US(r, j) != 0.0) { nzj = j; cnt++; }
        if (cnt == 1 && US(r, nzj) == 1.0 && freeRow[nzj] < 0) freeRow[nzj] = r;
    }
    for (int j = 0; j < dimker; j++) if (freeRow[j] < 0) return {};
    // S = [G2; Cfull]: only its rows that touch a transition DOF matter for the transition
    // solve; the normal matrix over the transition DOFs is small and dense
    if (nint > 4096) return {};
    struct SRow { const CsrView* A; int row; };
    std::vector<SRow> trows;
    for (const CsrView* A : {&G2, &Cf}) {
        // scanned in row chunks, each chunk's hits appended in chunk order: rows stay ascending
        constexpr int kCh = 64;
        const int rows = A->rows(), step = (rows + kCh - 1) / kCh;
        std::vector<std::vector<int>> hit(kCh);
#pragma omp parallel for schedule(dynamic, 1)
        for (int ch = 0; ch < kCh; ch++)
            for (int i = ch * step; i < std::min(rows, (ch + 1) * step); i++)
                for (int64_t q = A->ptr[i]; q < A->ptr[i + 1]; q++)
                    if (tpos[A->col[q]] >= 0) { hit[size_t(ch)].push_back(i); break; }
        for (const std::vector<int>& h : hit)
            for (int i : h) trows.push_back({A, i});
    }
    const size_t nt = size_t(std::max(nint, 1));
    std::vector<double> Mt(nt * nt, 0.0);  // sum over those rows of the outer product of transition entries
    for (const SRow& t : trows) {
        const CsrView& A = *t.A;
        for (int64_t p = A.ptr[t.row]; p < A.ptr[t.row + 1]; p++) {
            const int a = tpos[A.col[p]];
            if (a < 0) continue;
            for (int64_t q = A.ptr[t.row]; q < A.ptr[t.row + 1]; q++) {
                const int b = tpos[A.col[q]];
                if (b >= 0) Mt[size_t(a) * nt + size_t(b)] += A.val[p] * A.val[q];
            }
        }
    }
    std::vector<double> out(size_t(k) * n, 0.0);
    for (int c = 0; c < n; c++) {
        std::vector<double> y(size_t(nfull), 0.0);
        for (int i = 0; i < nfull; i++) if (tpos[i] < 0 && i % n == c) y[i] = 1.0;
        if (nint > 0) {
            // b = -GT^T (S y), over the rows that touch a transition DOF
            std::vector<double> b(static_cast<size_t>(nint), 0.0), x(static_cast<size_t>(nint), 0.0);
            for (const SRow& t : trows) {
                const CsrView& A = *t.A;
                double sy = 0.0;
                for (int64_t q = A.ptr[t.row]; q < A.ptr[t.row + 1]; q++) sy += A.val[q] * y[A.col[q]];
                for (int64_t q = A.ptr[t.row]; q < A.ptr[t.row + 1]; q++)
                    if (tpos[A.col[q]] >= 0) b[size_t(tpos[A.col[q]])] -= A.val[q] * sy;
            }
            // Jacobi-CG on the (possibly semidefinite, consistent) normal equations
            std::vector<double> r(b), z(static_cast<size_t>(nint)), p, q(static_cast<size_t>(nint));
            auto prec = [&](const std::vector<double>& a, std::vector<double>& o) {
                for (int i = 0; i < nint; i++) {
                    const double dgi = Mt[size_t(i) * nt + size_t(i)];
                    o[i] = dgi > 0.0 ? a[i] / dgi : 0.0;
                }
            };
            prec(r, z);
            p = z;
            double rz = 0.0, bb = 0.0;
            for (int i = 0; i < nint; i++) { rz += r[i] * z[i]; bb += b[i] * b[i]; }
            int it = 0;
            for (; it < 20000 && bb > 0.0; it++) {
                for (int i = 0; i < nint; i++) {
                    double sq = 0.0;
                    for (int j = 0; j < nint; j++) sq += Mt[size_t(i) * nt + size_t(j)] * p[j];
                    q[i] = sq;
                }
                double pq = 0.0;
                for (int i = 0; i < nint; i++) pq += p[i] * q[i];
                if (!(pq > 0.0)) break;
                const double al = rz / pq;
                double rr = 0.0;
                for (int i = 0; i < nint; i++) { x[i] += al * p[i]; r[i] -= al * q[i]; rr += r[i] * r[i]; }
                if (rr <= 1e-30 * bb) break;
                prec(r, z);
                double rz2 = 0.0;
                for (int i = 0; i < nint; i++) rz2 += r[i] * z[i];
                const double be = rz2 / rz;
                rz = rz2;
                for (int i = 0; i < nint; i++) p[i] = z[i] + be * p[i];
            }
            for (int i = 0; i < nfull; i++) if (tpos[i] >= 0) y[i] = x[tpos[i]];
            MK_LOG("[igm] near-null %d: %d transition DOFs, %zu rows, %d CG iterations, %.0f ms\n", c, nint,
                   trows.size(), it, now_ms() - t_nn);
        }
        double* xs = out.data() + size_t(c) * k;
        for (int i = 0; i < npn; i++) xs[i] = y[nonPart[i]];
        for (int j = 0; j < dimker; j++) xs[npn + j] = y[uniqueJ[freeRow[j]]];
        // null to round-off: ||X2F x||^2 against the same sum of |terms| (64 fixed row chunks)
        constexpr int kCh = 64;
        double pe[kCh], ps[kCh];
        const int rows = X2F.rows(), step = (rows + kCh - 1) / kCh;
#pragma omp parallel for schedule(static, 1)
        for (int ch = 0; ch < kCh; ch++) {
            double e = 0.0, m2 = 0.0;
            for (int i = ch * step; i < std::min(rows, (ch + 1) * step); i++) {
                double a = 0.0, m = 0.0;
                for (int64_t q = X2F.ptr[i]; q < X2F.ptr[i + 1]; q++) {
                    a += X2F.val[q] * xs[X2F.col[q]];
                    m += std::abs(X2F.val[q] * xs[X2F.col[q]]);
                }
                e += a * a;
                m2 += m * m;
            }
            pe[ch] = e;
            ps[ch] = m2;
        }
        double e = 0.0, s2 = 0.0;
        for (int ch = 0; ch < kCh; ch++) { e += pe[ch]; s2 += ps[ch]; }
        if (!(s2 > 0.0) || !(e <= 1e-20 * s2)) return {};
    }
    return out;
}


// ---- local injectivity (iter_rounding.h:33-113 with localInjectivity = true) ---------------
// Both LM passes run LMSolver with DiagonalDamping (levenberg_marquardt.h:13-83,140-231,
// lambda from 0.01) on the reference's objectives with the InjectiveBarrier rows
// (injective_barrier.h:10-110, iter_rounding_common.h:60-94).  Each direction solves the
// damped normal equations (J^T J + diag(lambda colsum J.^2)) d = -J^T E exactly as the
// reference's dampJ^T dampJ does, in this form:
//  * the 1e8 / 1e10 weighted rows (fixed values and their damping entries) are a low-rank
//    term W^T diag(omega) W handled by Sherman-Morrison-Woodbury;
//  * the rest, Hs, is assembled on the host from E = length X2F^T X2F plus the active barrier
//    rows and solved by multigrid PCG on the device (one hierarchy per direction);
//  * the naive pass eliminates f face by face (its f-block is block diagonal, 2N x 2N per
//    face) and solves the reduced system in x.
// Barrier values, derivatives, the gradients -J^T E and the energies are device kernels.

// barrier rows per (face i, pair j): r = (v_j x v_{j+1}) / V_ij, s = 0.5 (injective_barrier.h:53-103)
__global__ void k_barrier(int nF, int N, const double* field, const double* V, double* fB, double* der, double* S) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nF * N) return;
    const int i = t / N, j = t % N, j1 = (j + 1) % N;
    const double* a = field + size_t(2 * N) * i + 2 * j;
    const double* b = field + size_t(2 * N) * i + 2 * j1;
    const double v = V[t], s = 0.5;
    const double r = (a[0] * b[1] - a[1] * b[0]) / v;
    const double q = r / s;
    const double bar = q * q * q - 3.0 * q * q + 3.0 * q;
    double b2 = 1.0 / bar - 1.0;
    if (r <= 0.0) b2 = INFINITY;
    if (r >= s) b2 = 0.0;
    fB[t] = b2;
    if (!der) return;
    double d = 3.0 * (r * r / (s * s * s)) - 6.0 * (r / (s * s)) + 3.0 / s;
    if (r <= 0.0) d = INFINITY;
    if (r >= s) d = 0.0;
    double g = -d / (bar * bar);
    if (fabs(b2) < 10e-9) g = 0.0;
    else if (b2 == INFINITY) g = INFINITY;
    der[t] = g;
    S[4 * size_t(t) + 0] = b[1] / v;
    S[4 * size_t(t) + 1] = -b[0] / v;
    S[4 * size_t(t) + 2] = -a[1] / v;
    S[4 * size_t(t) + 3] = a[0] / v;
}
// out[c] = (S^T diag(der) e)[c] over the 2N nF field entries: column (face i, vector j,
// component k) is slot k of pair (i, j) and slot 2 + k of pair (i, j - 1)
__global__ void k_barrier_tmul(int nF, int N, const double* S, const double* der, const double* e, double* out) {
    const size_t c = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (c >= size_t(2 * N) * nF) return;
    const int i = int(c / (2 * N)), w = int(c % (2 * N)), j = w >> 1, k = w & 1;
    const size_t q0 = size_t(i) * N + j, q1 = size_t(i) * N + (j + N - 1) % N;
    out[c] = S[4 * q0 + k] * der[q0] * e[q0] + S[4 * q1 + 2 + k] * der[q1] * e[q1];
}
// per-face 2N x 2N operator: out = blocks[slot[face]] * in, or s * in for faces without a block
__global__ void k_face_apply(size_t l1, int N2, double s, const int* slot, const double* blocks, const double* in,
                             double* out) {
    const size_t c = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (c >= l1) return;
    const size_t f = c / N2;
    const int w = int(c % N2), sl = slot[f];
    if (sl < 0) { out[c] = s * in[c]; return; }
    const double* B = blocks + size_t(sl) * N2 * N2 + size_t(w) * N2;
    double a = 0.0;
    for (int b = 0; b < N2; b++) a += B[b] * in[f * N2 + b];
    out[c] = a;
}
// y = a x + b z + c w
__global__ void k_lin3(size_t n, double a, const double* x, double b, const double* z, double c, const double* w,
                       double* y) {
    const size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) y[i] = a * x[i] + b * z[i] + c * w[i];
}
// flag[f] = 1 where cut face f turns negative in (u, v) = functions 0, 1 of NF
__global__ void k_flips(int nF, int N, const int* idx, const double* NF, double* flag) {
    const int f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= nF) return;
    const double* a = NF + size_t(N) * idx[3 * f];
    const double* b = NF + size_t(N) * idx[3 * f + 1];
    const double* c = NF + size_t(N) * idx[3 * f + 2];
    const double d1x = b[0] - a[0], d1y = b[1] - a[1], d2x = c[0] - a[0], d2y = c[1] - a[1];
    flag[f] = (d1x * d2y - d1y * d2x < 0.0) ? 1.0 : 0.0;
}
// y[idx[t]] += val[t] (distinct indices)
__global__ void k_add_sparse(int m, const int* idx, const double* val, double* y) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < m) y[idx[t]] += val[t];
}

// small dense SPD inverse (Gauss-Jordan with partial pivoting), row-major n x n
std::vector<double> dense_inverse(std::vector<double> A, int n) {
    std::vector<double> I(size_t(n) * n, 0.0);
    for (int i = 0; i < n; i++) I[size_t(i) * n + i] = 1.0;
    for (int k = 0; k < n; k++) {
        int p = k;
        for (int i = k + 1; i < n; i++) if (std::abs(A[size_t(i) * n + k]) > std::abs(A[size_t(p) * n + k])) p = i;
        if (A[size_t(p) * n + k] == 0.0) throw Error(MK_ERR_SINGULAR, "injectivity: singular face block");
        if (p != k)
            for (int j = 0; j < n; j++) {
                std::swap(A[size_t(k) * n + j], A[size_t(p) * n + j]);
                std::swap(I[size_t(k) * n + j], I[size_t(p) * n + j]);
            }
        const double piv = A[size_t(k) * n + k];
        for (int j = 0; j < n; j++) { A[size_t(k) * n + j] /= piv; I[size_t(k) * n + j] /= piv; }
        for (int i = 0; i < n; i++) {
            if (i == k) continue;
            const double l = A[size_t(i) * n + k];
            if (l == 0.0) continue;
            for (int j = 0; j < n; j++) { A[size_t(i) * n + j] -= l * A[size_t(k) * n + j]; I[size_t(i) * n + j] -= l * I[size_t(k) * n + j]; }
        }
    }
    return I;
}

// H + diag(dv) for a matrix whose rows all hold their diagonal entry
void add_diagonal(SprsD& H, const std::vector<double>& dv) {
    for (int i = 0; i < H.rows(); i++) {
        bool found = false;
        for (int64_t q = H.ptr[i]; q < H.ptr[i + 1]; q++)
            if (H.col[q] == i) { H.val[q] += dv[size_t(i)]; found = true; break; }
        if (!found) throw Error(MK_ERR_SINGULAR, "injectivity: a variable no gradient row touches");
    }
}

// sparse sum of rows of M (row index list rows[], coefficients coef[]) as sorted (col, val)
void merge_rows(const SprsD& M, const std::vector<int>& rows, const std::vector<double>& coef,
                std::vector<std::pair<int, double>>& out) {
    out.clear();
    for (size_t a = 0; a < rows.size(); a++)
        for (int64_t q = M.ptr[rows[a]]; q < M.ptr[rows[a] + 1]; q++) out.emplace_back(M.col[q], coef[a] * M.val[q]);
    std::sort(out.begin(), out.end(), [](const auto& x, const auto& y) { return x.first < y.first; });
    size_t u = 0;
    for (size_t t = 0; t < out.size(); t++) {
        if (u && out[u - 1].first == out[t].first) out[u - 1].second += out[t].second;
        else out[u++] = out[t];
    }
    out.resize(u);
}

struct InjSolver {
    Ctx& K;
    const mk_op* A0;   // near-null vectors of the x-space operators
    int k;
    DevBuf<double> y, Z, drow, dd, dw, Bp, Xp, dwz;
    DevBuf<int> didx, dsel;
    bool jacobi = false;
    int jacobi_solves = 0;
    InjSolver(Ctx& K_, const mk_op* A0_, int k_) : K(K_), A0(A0_), k(k_), y(size_t(k_)), drow(size_t(k_)) {}

    // g += sum_i coef_i R_i (rows of R)
    void add_rows(const SprsD& R, const std::vector<double>& coef, double* g) {
        std::vector<int> rows(static_cast<size_t>(R.rows()));
        std::iota(rows.begin(), rows.end(), 0);
        std::vector<std::pair<int, double>> m;
        merge_rows(R, rows, coef, m);
        if (m.empty()) return;
        std::vector<int> ix(m.size());
        std::vector<double> vv(m.size());
        for (size_t t = 0; t < m.size(); t++) { ix[t] = m[t].first; vv[t] = m[t].second; }
        didx.upload(ix.data(), ix.size(), K.st());
        dw.upload(vv.data(), vv.size(), K.st());
        hipLaunchKernelGGL(k_add_sparse, dim3(nblk(m.size())), dim3(256), 0, K.st(), int(m.size()), didx.p, dw.p, g);
    }
    // R x on the device -> host
    std::vector<double> rowdots(const SprsD& R, const double* x) {
        std::vector<double> out(static_cast<size_t>(R.rows()), 0.0);
        if (!R.rows()) return out;
        RowSet rs;
        for (int i = 0; i < R.rows(); i++) rs.push(R, i);
        rs.sync(K.st());
        dd.ensure(size_t(R.rows()));
        K.rowdots(rs, x, dd.p);
        dd.download(out.data(), out.size(), K.st());
        K.sync();
        return out;
    }
    // multigrid PCG to the solve tolerance, at most 1000 iterations; an unconverged or broken-
    // down result is kept (an LM direction only has to lower the energy, which the caller
    // tests; a non-finite one makes the step a rejected one)
    void pcg(const mk_op* A, const double* b, double* x) {
        mk_solve_opts o2 = K.o;
        o2.maxiter = std::min(o2.maxiter, jacobi ? 5000 : 1000);
        if (jacobi) o2.precond = MK_PRECOND_JACOBI;
        const mk::SolveResult r = mk::solve_spd(K.c, A, b, x, o2);
        K.acc.iterations += r.iters;
        K.acc.refinements += r.refinements;
        K.acc.solves += 1;
        if (r.status == MK_ERR_ARG || r.status == MK_ERR_HIP || r.status == MK_ERR_ALLOC ||
            r.status == MK_ERR_UNSUPPORTED)
            throw Error(r.status, "IGM solve failed: " + std::string(mk_status_str(r.status)));
    }
    // (Hs + W^T diag(omega) W) d = g; Hs SPD (host CSR, k x k, from build()), W few sparse rows.
    // diag_dominant: the damping makes Hs well conditioned under Jacobi scaling (Hs holds
    // lambda diag(J^T J) on top of J^T J), so Jacobi PCG needs no hierarchy set-up; otherwise
    // a multigrid hierarchy of its own (Jacobi when aggregation cannot coarsen Hs).  key != 0:
    // Hs depends on key alone (lambda and the pass, no barrier row active), and the operator
    // with its hierarchy is kept for the next direction with the same key (up to 16).
    struct CachedOp {
        mk_op* op = nullptr;
        bool jacobi = false;
        ~CachedOp() { if (op) mk_op_destroy(op); }
    };
    std::map<uint64_t, std::unique_ptr<CachedOp>> cache;
    int cache_hits = 0;
    template <typename Build>
    void solve(Build&& build, uint64_t key, const SprsD& W, const std::vector<double>& omega, const double* g,
               double* d, bool diag_dominant) {
        std::unique_ptr<CachedOp> local;
        CachedOp* hc = nullptr;
        if (key) {
            const auto it = cache.find(key);
            if (it != cache.end()) { hc = it->second.get(); cache_hits++; }
        }
        if (!hc) {
            auto co = std::make_unique<CachedOp>();
            {
                SprsD Hs = build();
                co->op = mk::op_create_owned(K.c, k, std::move(Hs.ptr), std::move(Hs.col), std::move(Hs.val), K.o);
            }
            co->op->label = "igm_inj";
            co->op->nullvec = A0->nullvec;
            co->op->nnull = A0->nnull;
            co->jacobi = diag_dominant || K.o.precond != MK_PRECOND_AMG;
            if (!co->jacobi) {
                try {
                    mk::ensure_mg(K.c, co->op, K.o);
                } catch (const Error& e) {
                    if (e.code != MK_ERR_UNSUPPORTED) throw;
                    co->jacobi = true;
                }
            }
            hc = co.get();
            if (key && cache.size() < 16) cache.emplace(key, std::move(co));
            else local = std::move(co);
        }
        struct { mk_op* op; } h{hc->op};
        jacobi = hc->jacobi;
        if (jacobi && K.o.precond == MK_PRECOND_AMG) jacobi_solves++;
        hipLaunchKernelGGL(k_zero, dim3(nblk(size_t(k))), dim3(256), 0, K.st(), size_t(k), y.p);
        pcg(h.op, g, y.p);
        const int m = W.rows();
        if (m == 0) {
            MK_HIP(hipMemcpyAsync(d, y.p, size_t(k) * sizeof(double), hipMemcpyDeviceToDevice, K.st()));
            return;
        }
        Z.ensure(size_t(k) * size_t(m));
        RowSet ws;
        for (int i = 0; i < m; i++) ws.push(W, i);
        ws.sync(K.st());
        std::vector<double> M(static_cast<size_t>(m) * m), col(static_cast<size_t>(m));
        dd.ensure(size_t(m));
        // every precision mode batches the columns; MK_PREC_FP64 runs the V-cycle in FP64
        const bool blk = !jacobi && h.op->mg;
        if (blk) {
            // the Woodbury columns as blocks of up to 64 right-hand sides (one matrix read per step)
            const int mwmax = m > 32 ? 64 : m > 16 ? 32 : m > 8 ? 16 : 8;
            Bp.ensure(size_t(k) * mwmax);
            Xp.ensure(size_t(k) * mwmax);
            for (int b0 = 0, cnt = 0; b0 < m; b0 += cnt) {
                cnt = std::min(mwmax, m - b0);
                const int mw = cnt > 32 ? 64 : cnt > 16 ? 32 : cnt > 8 ? 16 : 8;
                std::vector<int> sel(static_cast<size_t>(mw), -1);
                for (int j = 0; j < cnt; j++) sel[size_t(j)] = b0 + j;
                dsel.upload(sel.data(), sel.size(), K.st());
                hipLaunchKernelGGL(k_zero, dim3(nblk(size_t(k) * mw)), dim3(256), 0, K.st(), size_t(k) * mw, Bp.p);
                hipLaunchKernelGGL(k_scatter_rows_packed, dim3(unsigned(mw)), dim3(64), 0, K.st(), mw, dsel.p, ws.ptr.p,
                                   ws.col.p, ws.val.p, Bp.p);
                const mk::BlockResult br = mk::block_pcg(K.c, h.op, h.op->mg, mw, Bp.p, Xp.p, K.o.tol,
                                                         std::min(K.o.maxiter, 1000), K.o.check_every,
                                                         K.o.precision == MK_PREC_FP64);
                K.acc.iterations += br.iters;
                K.acc.solves += cnt;
                if (br.status == MK_ERR_ARG || br.status == MK_ERR_HIP || br.status == MK_ERR_ALLOC ||
                    br.status == MK_ERR_UNSUPPORTED)
                    throw Error(br.status, "IGM block solve failed: " + std::string(mk_status_str(br.status)));
                hipLaunchKernelGGL(k_unpack_cols, dim3(nblk(size_t(k) * mw)), dim3(256), 0, K.st(), size_t(k), mw,
                                   dsel.p, (const double*)Xp.p, Z.p);
            }
        } else {
            for (int j = 0; j < m; j++) {
                double* z = Z.p + size_t(j) * k;
                hipLaunchKernelGGL(k_zero, dim3(nblk(size_t(k))), dim3(256), 0, K.st(), size_t(k), drow.p);
                const int64_t b = ws.hptr[size_t(j)], e = ws.hptr[size_t(j) + 1];
                if (e > b)
                    hipLaunchKernelGGL(k_scatter_row, dim3(nblk(size_t(e - b))), dim3(256), 0, K.st(), b, e, ws.col.p,
                                       ws.val.p, drow.p);
                hipLaunchKernelGGL(k_zero, dim3(nblk(size_t(k))), dim3(256), 0, K.st(), size_t(k), z);
                pcg(h.op, drow.p, z);
            }
        }
        // W Z (m x m), all pairs in one launch: WZ[i * m + j] = row i of W . Z[:, j]
        dwz.ensure(size_t(m) * m);
        hipLaunchKernelGGL(k_uz, dim3(nblk(size_t(m) * m)), dim3(256), 0, K.st(), m, size_t(k), ws.ptr.p, ws.col.p,
                           ws.val.p, (const double*)Z.p, dwz.p);
        dwz.download(M.data(), M.size(), K.st());
        K.sync();
        for (int i = 0; i < m; i++) M[size_t(i) * m + i] += 1.0 / omega[size_t(i)];
        K.rowdots(ws, y.p, dd.p);
        dd.download(col.data(), col.size(), K.st());
        K.sync();
        std::vector<double> cc = dense_solve(M, col, m);
        for (auto& v : cc) v = -v;
        std::vector<int> idx(static_cast<size_t>(m));
        std::iota(idx.begin(), idx.end(), 0);
        didx.upload(idx.data(), idx.size(), K.st());
        dw.upload(cc.data(), cc.size(), K.st());
        hipLaunchKernelGGL(k_combine, dim3(nblk(size_t(k))), dim3(256), 0, K.st(), size_t(k), m, y.p, Z.p,
                           (const int*)didx.p, dw.p, d);
    }
};

// LMSolver::solve with DiagonalDamping (lambda from 0.01).  energy(x) -> ||E(x)||^2;
// jac(x) evaluates the Jacobian state at x; dir(lambda, d) -> false when that Jacobian is not
// finite (the reference's direction is then NaN and no later step is taken); post(x) -> true
// ends the solve after an iteration (ST->post_iteration).
template <typename Energy, typename Jac, typename Dir, typename Post>
bool lm_damped(Ctx& K, size_t n, double* x, int maxIter, double funcTol, Energy&& energy, Jac&& jac, Dir&& dir,
               Post&& post, DevBuf<double>& d, DevBuf<double>& trial, int& iters) {
    d.ensure(n);
    trial.ensure(n);
    double lam = 0.01;
    jac(x);
    for (int it = 0;;) {
        iters++;
        if (!dir(lam, d.p)) return false;
        const double dn = std::sqrt(K.sqdiff(n, 1.0, d.p, nullptr));
        if (dn < funcTol) return true;
        const double e_prev = energy(x);
        hipLaunchKernelGGL(k_lin2, dim3(nblk(n)), dim3(256), 0, K.st(), n, 1.0, x, 1.0, d.p, trial.p);
        const double e_new = energy(trial.p);
        const bool acc = e_prev > e_new;
        if (acc) {
            MK_HIP(hipMemcpyAsync(x, trial.p, n * sizeof(double), hipMemcpyDeviceToDevice, K.st()));
            if (std::abs(e_prev - e_new) < funcTol) return false;
            jac(x);
        }
        lam = (acc && e_new != INFINITY) ? lam / 10.0 : lam * 10.0;
        if (post(x)) return true;
        it++;
        if (!(it <= maxIter)) return false;
    }
}

struct InjInputs {
    int k, nfull, l1, N, nF;
    double length;
    const SprsD* UF;
    const SprsD* X2F;    // G2 UFull (without length)
    const SprsD* E;      // length X2F^T X2F
    const SprsD* Cm;     // UFull rows at fixedIdcs
    const double* fixedVals;
    const double* F2;
    const std::vector<double>* XSmall;
    const mk_op* A0;
    mk_op* B;            // length X2F (device)
    mk_op* Bt;           // length X2F^T (device)
};

// barrier state of one field (device), with host copies of der and S after a Jacobian update
struct BarrierState {
    DevBuf<double> V, fB, der, S, tmul;
    std::vector<double> hder, hS;
    int N, nF;
    void init(Ctx& K, const double* F2, int N_, int nF_) {
        N = N_;
        nF = nF_;
        std::vector<double> v(size_t(N) * nF);
        for (int i = 0; i < nF; i++)
            for (int j = 0; j < N; j++) {
                const double* a = F2 + size_t(2 * N) * i + 2 * j;
                const double* b = F2 + size_t(2 * N) * i + 2 * ((j + 1) % N);
                v[size_t(i) * N + j] = std::sqrt(a[0] * a[0] + a[1] * a[1]) * std::sqrt(b[0] * b[0] + b[1] * b[1]);
            }
        V.upload(v.data(), v.size(), K.st());
        fB.alloc(v.size());
        der.alloc(v.size());
        S.alloc(4 * v.size());
        tmul.alloc(size_t(2 * N) * nF);
    }
    // ||fB(field)||^2
    double energy(Ctx& K, const double* field) {
        const size_t m = size_t(N) * nF;
        hipLaunchKernelGGL(k_barrier, dim3(nblk(m)), dim3(256), 0, K.st(), nF, N, field, V.p, fB.p, nullptr, nullptr);
        return K.sqdiff(m, 1.0, fB.p, nullptr);
    }
    // fB, der, S at field; tmul = S^T diag(der) fB; false when a derivative is not finite
    bool update(Ctx& K, const double* field) {
        const size_t m = size_t(N) * nF;
        hipLaunchKernelGGL(k_barrier, dim3(nblk(m)), dim3(256), 0, K.st(), nF, N, field, V.p, fB.p, der.p, S.p);
        hipLaunchKernelGGL(k_barrier_tmul, dim3(nblk(2 * m)), dim3(256), 0, K.st(), nF, N, S.p, der.p, fB.p, tmul.p);
        hder.resize(m);
        hS.resize(4 * m);
        der.download(hder.data(), m, K.st());
        S.download(hS.data(), 4 * m, K.st());
        K.sync();
        for (double v : hder) if (!std::isfinite(v)) return false;
        return true;
    }
};

// Hs for the seamless pass: wI^2 length E + wC^2 I + sum_q wB^2 v_q^T v_q plus the damping
// diagonal lambda (wI^2 length diag E + wC^2 + wB^2 sum_q v_q.^2); v_q = der_q S_q X2F'.
// Barrier rows with wB^2 |v_q|^2 above tau (near-flipped pairs: der grows like 1/r^2) go to
// `stiff` instead, for the low-rank part of the solve.
using SparseRow = std::vector<std::pair<int, double>>;
SprsD seamless_soft(const InjInputs& in, const BarrierState& bs, double lam, double wI, double wC, double wB,
                    std::vector<SparseRow>& stiff) {
    const int k = in.k;
    SprsD H = *in.E;   // X2F'^T X2F' = length E
    for (auto& v : H.val) v *= wI * wI * in.length;
    std::vector<double> dv(static_cast<size_t>(k), 0.0);
    double dmax = 0.0;
    for (int i = 0; i < k; i++)
        for (int64_t q = in.E->ptr[i]; q < in.E->ptr[i + 1]; q++)
            if (in.E->col[q] == i) {
                dv[size_t(i)] = lam * wI * wI * in.length * in.E->val[q];
                dmax = std::max(dmax, wI * wI * in.length * in.E->val[q]);
            }
    for (auto& v : dv) v += wC * wC * (1.0 + lam);
    const double tau = 1e4 * (dmax + wC * wC);
    stiff.clear();
    std::vector<TripD> T;
    SparseRow vq;
    const int N = bs.N;
    for (int i = 0; i < bs.nF; i++)
        for (int j = 0; j < N; j++) {
            const size_t q = size_t(i) * N + j;
            const double g = bs.hder[q];
            if (g == 0.0) continue;
            const int j1 = (j + 1) % N;
            const int r0 = 2 * N * i;
            std::vector<int> rows{r0 + 2 * j, r0 + 2 * j + 1, r0 + 2 * j1, r0 + 2 * j1 + 1};
            std::vector<double> cf(4);
            for (int s = 0; s < 4; s++) cf[size_t(s)] = g * bs.hS[4 * q + s] * in.length;
            merge_rows(*in.X2F, rows, cf, vq);
            double n2 = 0.0;
            for (auto& a : vq) {
                dv[size_t(a.first)] += lam * wB * wB * a.second * a.second;
                n2 += a.second * a.second;
            }
            if (wB * wB * n2 > tau) { stiff.push_back(vq); continue; }
            for (auto& a : vq)
                for (auto& b : vq) T.emplace_back(a.first, b.first, wB * wB * a.second * b.second);
        }
    if (!T.empty()) {
        SprsD Bar(k, k);
        Bar.setFromTriplets(T);
        H = H + Bar;
    }
    add_diagonal(H, dv);
    return H;
}

// W with the rows `extra` appended (weight w each)
void append_rows(SprsD& W, std::vector<double>& omega, const std::vector<SparseRow>& extra, double w) {
    if (extra.empty()) return;
    std::vector<TripD> T;
    for (int i = 0; i < W.rows(); i++)
        for (int64_t q = W.ptr[i]; q < W.ptr[i + 1]; q++) T.emplace_back(i, W.col[q], W.val[q]);
    int r = W.rows();
    for (const auto& row : extra) {
        for (const auto& e : row) T.emplace_back(r, e.first, e.second);
        omega.push_back(w);
        r++;
    }
    SprsD W2(r, W.cols());
    W2.setFromTriplets(T);
    W = std::move(W2);
}

// the low-rank stiff rows: R (weight wR each) and unit rows on the columns R touches
// (weight lambda wD colsum(R.^2), the damping entries of rows weighted wD)
void stiff_rows(const SprsD& R, double wR, double wD, double lam, int k, SprsD& W, std::vector<double>& omega) {
    std::vector<double> cs(size_t(k), 0.0);
    std::vector<int> cols;
    for (int64_t q = 0; q < R.nonZeros(); q++) {
        if (cs[size_t(R.col[q])] == 0.0 && R.val[q] != 0.0) cols.push_back(R.col[q]);
        cs[size_t(R.col[q])] += R.val[q] * R.val[q];
    }
    std::sort(cols.begin(), cols.end());
    cols.erase(std::unique(cols.begin(), cols.end()), cols.end());
    std::vector<TripD> T;
    omega.clear();
    for (int i = 0; i < R.rows(); i++) {
        for (int64_t q = R.ptr[i]; q < R.ptr[i + 1]; q++) T.emplace_back(i, R.col[q], R.val[q]);
        omega.push_back(wR);
    }
    int r = R.rows();
    for (int c : cols) {
        const double om = lam * wD * cs[size_t(c)];
        if (!(om > 0.0)) continue;
        T.emplace_back(r++, c, 1.0);
        omega.push_back(om);
    }
    W = SprsD(r, k);
    W.setFromTriplets(T);
}

// LMSolver over z = (x, f) for the naive pass (iter_rounding_init.h:60-72, weights
// iter_rounding_init.h:184-191: wInteg = wConst = 1e4, wClose = 1, wBarrier = 1e-4).
// J = [wI (-B | I); wC (0 | I); wK (Cm | 0); wI wC D S (0 | I)], B = length X2F.  Its normal
// matrix has the face-block-diagonal f-block Hff = a I + K (a = (1 + lambda)(wI^2 + wC^2),
// K = (wI wC)^2 (S^T D^2 S + lambda diag colsum (DS).^2)); with P = Hff^-1 and
// Q = P ((a - wI^2) I + K) = I - wI^2 P the reduced system is
//   (wI^2 B^T Q B + wK^2 Cm^T Cm + lambda diag_x) dx = wI^2 B^T (Q res + P gb) - wK^2 Cm^T cres,
//   df = P (-wI^2 (f - B (x + dx)) + gb),
// res = f - B x, gb = -wC^2 (f - F2) - wI wC wB S^T D fB, cres = Cm x - fixedVals (Q is formed
// without the cancellation of I - wI^2 P).
struct NaivePass {
    Ctx& K;
    const InjInputs& in;
    InjSolver& sol;
    BarrierState& bs;
    const double wI = 10e3, wK = 10e3, wC = 1.0, wB = 1e-4;
    DevBuf<double> dF2, Bx, res, gb, t1, t2, tk, gx, dP, dQ;
    DevBuf<int> dslot;
    std::vector<double> cres;
    const double* zcur = nullptr;
    bool finite = true;
    double mean_ediag = 0.0;   // mean diagonal entry of E
    NaivePass(Ctx& K_, const InjInputs& in_, InjSolver& sol_, BarrierState& bs_) : K(K_), in(in_), sol(sol_), bs(bs_) {
        const size_t l1 = size_t(in.l1), k = size_t(in.k);
        dF2.upload(in.F2, l1, K.st());
        Bx.alloc(l1); res.alloc(l1); gb.alloc(l1); t1.alloc(l1); t2.alloc(l1);
        tk.alloc(k); gx.alloc(k);
        dslot.alloc(size_t(in.nF));
        for (int i = 0; i < in.k; i++)
            for (int64_t t = in.E->ptr[i]; t < in.E->ptr[i + 1]; t++)
                if (in.E->col[t] == i) mean_ediag += in.E->val[t];
        mean_ediag /= std::max(1, in.k);
    }
    std::vector<double> cm_res(const double* x) {
        std::vector<double> r = sol.rowdots(*in.Cm, x);
        for (int i = 0; i < in.Cm->rows(); i++) r[size_t(i)] -= in.fixedVals[i];
        return r;
    }
    double energy(const double* z) {
        const size_t l1 = size_t(in.l1);
        const double* f = z + in.k;
        mk::spmv64(K.c, in.B, z, Bx.p, nullptr, mk::SPMV_PLAIN, nullptr, nullptr);
        const double a = K.sqdiff(l1, 1.0, f, Bx.p);
        const double b = K.sqdiff(l1, 1.0, f, dF2.p);
        double c = 0.0;
        for (double v : cm_res(z)) c += v * v;
        const double d = bs.energy(K, f);
        return wI * wI * a + wC * wC * b + wK * wK * c + wB * wB * d;
    }
    void jac(const double* z) {
        const size_t l1 = size_t(in.l1);
        const double* f = z + in.k;
        zcur = z;
        mk::spmv64(K.c, in.B, z, Bx.p, nullptr, mk::SPMV_PLAIN, nullptr, nullptr);
        finite = bs.update(K, f);
        hipLaunchKernelGGL(k_lin2, dim3(nblk(l1)), dim3(256), 0, K.st(), l1, 1.0, f, -1.0, Bx.p, res.p);
        hipLaunchKernelGGL(k_lin3, dim3(nblk(l1)), dim3(256), 0, K.st(), l1, -wC * wC, f, wC * wC, dF2.p,
                           -wI * wC * wB, bs.tmul.p, gb.p);
        cres = cm_res(z);
    }
    bool dir(double lam, double* dz) {
        if (!finite) return false;
        const int N = in.N, N2 = 2 * N, nF = in.nF, k = in.k;
        const size_t l1 = size_t(in.l1);
        const double a = (1.0 + lam) * (wI * wI + wC * wC);
        const double am = lam * wI * wI + (1.0 + lam) * wC * wC;   // a - wI^2
        const double p = 1.0 / a, q = am / a;
        const double wf = wI * wC;
        std::vector<int> slot(size_t(nF), -1);
        std::vector<double> Pb, Qb;
        std::vector<int> act;
        for (int i = 0; i < nF; i++) {
            bool any = false;
            for (int j = 0; j < N; j++) any |= bs.hder[size_t(i) * N + j] != 0.0;
            if (!any) continue;
            slot[size_t(i)] = int(act.size());
            act.push_back(i);
            std::vector<double> Kb(size_t(N2) * N2, 0.0), cs(size_t(N2), 0.0);
            for (int j = 0; j < N; j++) {
                const size_t qq = size_t(i) * N + j;
                const double g = bs.hder[qq];
                if (g == 0.0) continue;
                const int cl[4] = {2 * j, 2 * j + 1, 2 * ((j + 1) % N), 2 * ((j + 1) % N) + 1};
                double sv[4];
                for (int s = 0; s < 4; s++) sv[s] = g * bs.hS[4 * qq + s];
                for (int s = 0; s < 4; s++) {
                    cs[size_t(cl[s])] += sv[s] * sv[s];
                    for (int t = 0; t < 4; t++) Kb[size_t(cl[s]) * N2 + cl[t]] += sv[s] * sv[t];
                }
            }
            for (int c = 0; c < N2; c++) Kb[size_t(c) * N2 + c] += lam * cs[size_t(c)];
            for (auto& v : Kb) v *= wf * wf;
            std::vector<double> H(Kb);
            for (int c = 0; c < N2; c++) H[size_t(c) * N2 + c] += a;
            std::vector<double> P = dense_inverse(H, N2);
            std::vector<double> Qi(size_t(N2) * N2, 0.0);   // P (am I + K)
            for (int r = 0; r < N2; r++)
                for (int c = 0; c < N2; c++) {
                    double s = am * P[size_t(r) * N2 + c];
                    for (int t = 0; t < N2; t++) s += P[size_t(r) * N2 + t] * Kb[size_t(t) * N2 + c];
                    Qi[size_t(r) * N2 + c] = s;
                }
            Pb.insert(Pb.end(), P.begin(), P.end());
            Qb.insert(Qb.end(), Qi.begin(), Qi.end());
        }
        dslot.upload(slot.data(), slot.size(), K.st());
        if (!act.empty()) {
            dP.upload(Pb.data(), Pb.size(), K.st());
            dQ.upload(Qb.data(), Qb.size(), K.st());
        }
        const double* Pp = act.empty() ? nullptr : dP.p;
        const double* Qp = act.empty() ? nullptr : dQ.p;
        // reduced right-hand side
        hipLaunchKernelGGL(k_face_apply, dim3(nblk(l1)), dim3(256), 0, K.st(), l1, N2, q, dslot.p, Qp, res.p, t1.p);
        hipLaunchKernelGGL(k_face_apply, dim3(nblk(l1)), dim3(256), 0, K.st(), l1, N2, p, dslot.p, Pp, gb.p, t2.p);
        hipLaunchKernelGGL(k_lin2, dim3(nblk(l1)), dim3(256), 0, K.st(), l1, wI * wI, t1.p, wI * wI, t2.p, t1.p);
        mk::spmv64(K.c, in.Bt, t1.p, gx.p, nullptr, mk::SPMV_PLAIN, nullptr, nullptr);
        {
            std::vector<double> cf(cres.size());
            for (size_t i = 0; i < cres.size(); i++) cf[i] = -wK * wK * cres[i];
            sol.add_rows(*in.Cm, cf, gx.p);
        }
        // reduced operator: wI^2 q length E + lambda wI^2 length diag E + active face terms + pin
        // pin weight of the fixed rows inside Hs (the rest, wK^2 - pin, goes to the Woodbury
        // rows): the mean diagonal of the Poisson part, a function of lambda alone
        const double L = in.length;
        double dmean = std::min(wI * wI * L * (q + lam) * mean_ediag, 0.5 * wK * wK);
        if (!(dmean > 0.0)) dmean = 1.0;
        auto build = [&]() {
            SprsD H = *in.E;
            for (auto& v : H.val) v *= wI * wI * q * L;
            std::vector<double> dv(static_cast<size_t>(k), 0.0);
            for (int i = 0; i < k; i++)
                for (int64_t t = in.E->ptr[i]; t < in.E->ptr[i + 1]; t++)
                    if (in.E->col[t] == i) dv[size_t(i)] = lam * wI * wI * L * in.E->val[t];
            if (!act.empty()) {
                std::vector<TripD> T;
                const SprsD& X = *in.X2F;
                for (size_t s = 0; s < act.size(); s++) {
                    const int i = act[s];
                    const double* Qi = Qb.data() + s * size_t(N2) * N2;
                    for (int ra = 0; ra < N2; ra++)
                        for (int rb = 0; rb < N2; rb++) {
                            const double cab = wI * wI * (Qi[size_t(ra) * N2 + rb] - (ra == rb ? q : 0.0)) * L * L;
                            if (cab == 0.0) continue;
                            const int A = N2 * i + ra, Bq = N2 * i + rb;
                            for (int64_t u = X.ptr[A]; u < X.ptr[A + 1]; u++)
                                for (int64_t v = X.ptr[Bq]; v < X.ptr[Bq + 1]; v++)
                                    T.emplace_back(X.col[u], X.col[v], cab * X.val[u] * X.val[v]);
                        }
                }
                SprsD Bar(k, k);
                Bar.setFromTriplets(T);
                H = H + Bar;
            }
            add_diagonal(H, dv);
            if (in.Cm->rows()) H = H + (in.Cm->transpose() * (*in.Cm)).scaled(dmean);
            return H;
        };
        SprsD W;
        std::vector<double> omega;
        stiff_rows(*in.Cm, wK * wK - dmean, wK * wK, lam, k, W, omega);
        // Hs = lambda wI^2 length (E / (1 + lambda) + diag E) + ...: Jacobi-conditioned for
        // every lambda well above wC^2 / wI^2 = 1e-8; without active faces it depends on lambda
        // alone and is cached
        uint64_t key = 0;
        if (act.empty()) { std::memcpy(&key, &lam, sizeof(key)); key ^= 0x5a5a5a5a00000000ull; }
        sol.solve(build, key, W, omega, gx.p, dz, lam >= 1e-6);
        // df = P (-wI^2 (f - B (x + dx)) + gb)
        hipLaunchKernelGGL(k_lin2, dim3(nblk(size_t(k))), dim3(256), 0, K.st(), size_t(k), 1.0, zcur, 1.0, dz, tk.p);
        mk::spmv64(K.c, in.B, tk.p, t2.p, nullptr, mk::SPMV_PLAIN, nullptr, nullptr);
        hipLaunchKernelGGL(k_lin3, dim3(nblk(l1)), dim3(256), 0, K.st(), l1, -wI * wI, zcur + k, wI * wI, t2.p, 1.0,
                           gb.p, t2.p);
        hipLaunchKernelGGL(k_face_apply, dim3(nblk(l1)), dim3(256), 0, K.st(), l1, N2, p, dslot.p, Pp, t2.p, dz + k);
        return true;
    }
};

// LMSolver over x for one seamless round (iter_rounding_loop.h:77-86, weights
// iter_rounding_loop.h:109: wInteg = 1, wConst = 1e5, wClose = 0.01, wBarrier = 1e-4):
// J = [wI X2F'; wC I; wK U; wB D S X2F'], X2F' = length X2F.
struct SeamlessPass {
    Ctx& K;
    const InjInputs& in;
    InjSolver& sol;
    BarrierState& bs;
    const double wI = 1.0, wK = 10e4, wC = 0.01, wB = 1e-4;
    const double* F2_;     // device, l1
    const double* xPrev;   // device, k
    SprsD U;
    std::vector<double> r;
    DevBuf<double> field, t1, g;
    std::vector<double> ures;
    bool finite = true;
    SeamlessPass(Ctx& K_, const InjInputs& in_, InjSolver& sol_, BarrierState& bs_, const double* F2u,
                 const double* xp)
        : K(K_), in(in_), sol(sol_), bs(bs_), F2_(F2u), xPrev(xp) {
        field.alloc(size_t(in.l1));
        t1.alloc(size_t(in.l1));
        g.alloc(size_t(in.k));
    }
    std::vector<double> u_res(const double* x) {
        std::vector<double> v = sol.rowdots(U, x);
        for (size_t i = 0; i < v.size(); i++) v[i] -= r[i];
        return v;
    }
    double energy(const double* x) {
        const size_t l1 = size_t(in.l1), k = size_t(in.k);
        mk::spmv64(K.c, in.B, x, field.p, nullptr, mk::SPMV_PLAIN, nullptr, nullptr);
        const double a = K.sqdiff(l1, 1.0, field.p, F2_);
        const double b = K.sqdiff(k, 1.0, x, xPrev);
        double c = 0.0;
        for (double v : u_res(x)) c += v * v;
        const double d = bs.energy(K, field.p);
        return wI * wI * a + wC * wC * b + wK * wK * c + wB * wB * d;
    }
    void jac(const double* x) {
        const size_t l1 = size_t(in.l1), k = size_t(in.k);
        mk::spmv64(K.c, in.B, x, field.p, nullptr, mk::SPMV_PLAIN, nullptr, nullptr);
        finite = bs.update(K, field.p);
        // -J^T E = X2F'^T (wI^2 (F2_ - field) - wB^2 S^T D fB) - wC^2 (x - xPrev) - wK^2 U^T (U x - r)
        hipLaunchKernelGGL(k_lin3, dim3(nblk(l1)), dim3(256), 0, K.st(), l1, wI * wI, F2_, -wI * wI, field.p,
                           -wB * wB, bs.tmul.p, t1.p);
        mk::spmv64(K.c, in.Bt, t1.p, g.p, nullptr, mk::SPMV_PLAIN, nullptr, nullptr);
        hipLaunchKernelGGL(k_lin3, dim3(nblk(k)), dim3(256), 0, K.st(), k, 1.0, g.p, -wC * wC, x, wC * wC, xPrev, g.p);
        ures = u_res(x);
        std::vector<double> cf(ures.size());
        for (size_t i = 0; i < ures.size(); i++) cf[i] = -wK * wK * ures[i];
        sol.add_rows(U, cf, g.p);
    }
    bool dir(double lam, double* d) {
        if (!finite) return false;
        bool active = false;
        for (double v : bs.hder) active |= v != 0.0;
        // with barrier rows active Hs (and the stiff barrier rows) are assembled now; without,
        // Hs depends on lambda alone and is assembled only when no cached operator has it
        std::vector<SparseRow> stiff;
        SprsD H;
        if (active) H = seamless_soft(in, bs, lam, wI, wC, wB, stiff);
        SprsD W;
        std::vector<double> omega;
        stiff_rows(U, wK * wK, wK * wK, lam, in.k, W, omega);
        append_rows(W, omega, stiff, wB * wB);
        uint64_t key = 0;
        if (!active) std::memcpy(&key, &lam, sizeof(key));
        // Hs = A0 + lambda diag(A0) + ...: Jacobi condition about 2 / lambda.  Below
        // lambda = 0.1 multigrid with block PCG for the Woodbury columns; above it (rejected
        // steps) Jacobi, where aggregation of the diagonal-heavy Hs stalls
        sol.solve([&]() {
            if (active) return std::move(H);
            std::vector<SparseRow> none;
            return seamless_soft(in, bs, lam, wI, wC, wB, none);
        }, key, W, omega, g.p, d, lam >= 0.1);
        return true;
    }
};

// iterative_rounding with localInjectivity = true: the naive LM pass, then (seamless) the
// rounding loop of SeamlessIntegration with a damped barrier LM per round
// (iter_rounding.h:82-112, iter_rounding_loop.h:29-100).
template <typename Record>
void injective_rounding(Ctx& K, const InjInputs& in, const mk_igm_flags* flags, const int32_t* singularIdcs,
                        int nsing, const int32_t* integerIdcs, int nint, double* fullx, int* success, Record&& record,
                        mk_igm_stats& S) {
    const int k = in.k;
    const size_t l1 = size_t(in.l1);
    InjSolver sol(K, in.A0, k);
    BarrierState bs;
    bs.init(K, in.F2, in.N, in.nF);
    DevBuf<double> z(size_t(k) + l1), dd, trial;
    z.upload(in.XSmall->data(), size_t(k), K.st());
    MK_HIP(hipMemcpyAsync(z.p + k, in.F2, l1 * sizeof(double), hipMemcpyHostToDevice, K.st()));
    double t0 = now_ms();
    int lm_iters = 0;
    {
        // the naive pass's iter_cb: locally_injectivity_check (parameterization_integ.ipp:45-63)
        // when the caller set one (mk_igm_injectivity_check)
        const mk_inj_check_data& ck = K.c->inj_check;
        const bool checked = ck.nF > 0 && ck.cols == in.nfull;
        OpHolder hN;
        DevBuf<int> didx;
        DevBuf<double> dNF, dflag;
        if (checked) {
            SprsD f2n(ck.rows, ck.cols);
            f2n.ptr.assign(ck.ptr.begin(), ck.ptr.end());
            f2n.col.assign(ck.col.begin(), ck.col.end());
            f2n.val.assign(ck.val.begin(), ck.val.end());
            const SprsD Nf = f2n * (*in.UF);
            hN.op = make_rect_op(K.c, Nf, K.o);
            hN.op->label = "igm_inj_check";
            didx.upload(ck.idx.data(), ck.idx.size(), K.st());
            dNF.alloc(size_t(ck.rows));
            dflag.alloc(size_t(ck.nF));
        }
        // a host callback (mk_set_naive_iter_callback) is the iter_cb when set: it gets
        // fullx = UFull x of the iterate's x-block
        OpHolder hU;
        DevBuf<double> dfull;
        std::vector<double> hfull;
        if (K.c->naive_cb) {
            hU.op = make_rect_op(K.c, *in.UF, K.o);
            hU.op->label = "igm_iter_cb";
            dfull.alloc(size_t(in.nfull));
            hfull.resize(size_t(in.nfull));
        }
        auto post = [&](const double* v) {
            if (K.c->naive_cb) {
                mk::spmv64(K.c, hU.op, v, dfull.p, nullptr, mk::SPMV_PLAIN, nullptr, nullptr);
                dfull.download(hfull.data(), hfull.size(), K.st());
                K.sync();
                return K.c->naive_cb(K.c->naive_user, hfull.data(), in.nfull) != 0;
            }
            if (!checked) return false;
            mk::spmv64(K.c, hN.op, v, dNF.p, nullptr, mk::SPMV_PLAIN, nullptr, nullptr);
            hipLaunchKernelGGL(k_flips, dim3(nblk(size_t(ck.nF))), dim3(256), 0, K.st(), ck.nF, ck.N, didx.p, dNF.p,
                               dflag.p);
            return K.sqdiff(size_t(ck.nF), 1.0, dflag.p, nullptr) == 0.0;
        };
        NaivePass np(K, in, sol, bs);
        lm_damped(K, size_t(k) + l1, z.p, 100, 10e-10, [&](const double* v) { return np.energy(v); },
                  [&](const double* v) { np.jac(v); }, [&](double lam, double* d) { return np.dir(lam, d); }, post,
                  dd, trial, lm_iters);
    }
    S.naive_ms = now_ms() - t0;
    MK_LOG("[igm] injective naive LM %.0f ms, %d iterations\n", S.naive_ms, lm_iters);
    sol.cache.clear();   // the kept operators are the naive pass's; the rounding passes key their own
    std::vector<double> xs(static_cast<size_t>(k));
    z.download(xs.data(), xs.size(), K.st());
    K.sync();
    *success = 1;
    if (!flags->seamless) {
        in.UF->multiply(xs.data(), fullx);
        return;
    }
    t0 = now_ms();
    const SprsD& UF = *in.UF;
    DevBuf<double> xcur(static_cast<size_t>(k)), xprev(static_cast<size_t>(k)), dcand;
    MK_HIP(hipMemcpyAsync(xcur.p, z.p, size_t(k) * sizeof(double), hipMemcpyDeviceToDevice, K.st()));
    const double* F2u = z.p + k;   // F2_ = f of the naive pass
    std::vector<int> singular(singularIdcs, singularIdcs + nsing), integer(integerIdcs, integerIdcs + nint);
    RowSet cand;
    std::vector<int> candPos(size_t(in.nfull), -1), candIdx;
    for (int i : singular) if (candPos[i] < 0) { candPos[i] = int(candIdx.size()); candIdx.push_back(i); cand.push(UF, i); }
    for (int i : integer) if (candPos[i] < 0) { candPos[i] = int(candIdx.size()); candIdx.push_back(i); cand.push(UF, i); }
    cand.sync(K.st());
    dcand.alloc(std::max<size_t>(candIdx.size(), 1));
    std::vector<int> left = flags->round_seams ? integer : singular;
    std::vector<int> fixedI;
    std::vector<double> fixedV;
    std::vector<double> Xc(candIdx.size());
    bool roundedSingulars = false, rounded = false;
    int rounds = 0, lm_rounds = 0;
    while (!left.empty()) {
        MK_HIP(hipMemcpyAsync(xprev.p, xcur.p, size_t(k) * sizeof(double), hipMemcpyDeviceToDevice, K.st()));
        K.rowdots(cand, xcur.p, dcand.p);
        dcand.download(Xc.data(), Xc.size(), K.st());
        K.sync();
        double minDif = 3276700.0;
        int minIdx = -1;
        for (size_t i = 0; i < left.size(); i++) {
            const double v = Xc[size_t(candPos[left[i]])];
            const double dif = std::fabs(v - std::round(v));
            if (dif < minDif) { minIdx = int(i); minDif = dif; }
        }
        const int pick = left[size_t(minIdx)];
        const double Xp = Xc[size_t(candPos[pick])];
        fixedI.push_back(pick);
        fixedV.push_back(double(int(std::round(Xp))));
        record(pick, int(std::round(Xp)), minDif);
        left.erase(left.begin() + minIdx);
        if (left.empty() && !flags->round_seams && !roundedSingulars) {
            left = integer;
            roundedSingulars = true;
        }
        rounds++;
        if (!(minDif > 10e-7)) continue;
        rounded = true;
        lm_rounds++;
        SeamlessPass sp(K, in, sol, bs, F2u, xprev.p);
        sp.U = UF.rows_of(fixedI);
        sp.r = fixedV;
        int it = 0;
        lm_damped(K, size_t(k), xcur.p, 100, 1e-7, [&](const double* v) { return sp.energy(v); },
                  [&](const double* v) { sp.jac(v); }, [&](double lam, double* d) { return sp.dir(lam, d); },
                  [](const double*) { return false; }, dd, trial, it);
        lm_iters += it;
        if (lm_rounds % 5 == 0) MK_LOG("[igm] injective rounding: %d rounds, %d LM iterations\n", rounds, lm_iters);
        // post_checking (iter_rounding_loop.h:89-100)
        std::vector<double> ux = sol.rowdots(sp.U, xcur.p);
        double mx = -1.0;
        for (double v : ux) mx = std::max(mx, std::abs(v - std::round(v)));
        if (!(mx <= 10e-7)) { *success = 0; break; }
    }
    S.rounds = rounds;
    S.lm_rounds = lm_rounds;
    S.n_fixed = int(fixedI.size());
    MK_LOG("[igm] injective rounding %.0f ms: %d rounds, %d with LM, %d LM iterations (%d directions by Jacobi "
           "PCG, %d on a cached operator)\n", now_ms() - t0, rounds, lm_rounds, lm_iters, sol.jacobi_solves,
           sol.cache_hits);
    if (rounded) {
        xcur.download(xs.data(), xs.size(), K.st());
        K.sync();
        UF.multiply(xs.data(), fullx);
    } else {
        std::copy(in.XSmall->begin(), in.XSmall->end(), xs.begin());
        UF.multiply(xs.data(), fullx);
    }
    S.rounding_ms = now_ms() - t0;
}
}  // namespace

extern "C" mk_status mk_iterative_rounding(
    mk_ctx* c, int nfixed, const int32_t* fixedIdcs, const double* fixedVals, int nsing, const int32_t* singularIdcs,
    int nint, const int32_t* integerIdcs, double length, int c_rows, int c_cols, const int64_t* c_ptr,
    const int32_t* c_col, const double* c_val, int g_rows, int g_cols, const int64_t* g_ptr, const int32_t* g_col,
    const double* g_val, int N, int n, int nF, const mk_igm_flags* flags, const double* F2, const mk_solve_opts* o_,
    double* fullx, int* success, int max_dec, int32_t* dec_index, int32_t* dec_value, double* dec_margin, int* ndec,
    mk_igm_stats* stats) {
    if (!c || !flags || !F2 || !fullx || !success || !g_ptr || g_rows != 2 * N * nF || c_cols != g_cols) return MK_ERR_ARG;
    try {
        mk_solve_opts o;
        mk_default_opts(&o);
        if (o_) o = *o_;
        MK_HIP(hipSetDevice(c->device));
        mk::stats_reset(c);
        Ctx K(c, o);
        mk_igm_stats S{};
        const double t_all = now_ms();
        const int nfull = g_cols;
        for (int i = 0; i < nfixed; i++) MK_CHECK(fixedIdcs[i] >= 0 && fixedIdcs[i] < nfull, MK_ERR_ARG, "fixedIdcs");
        for (int i = 0; i < nsing; i++) MK_CHECK(singularIdcs[i] >= 0 && singularIdcs[i] < nfull, MK_ERR_ARG, "singularIdcs");
        for (int i = 0; i < nint; i++) MK_CHECK(integerIdcs[i] >= 0 && integerIdcs[i] < nfull, MK_ERR_ARG, "integerIdcs");
        // ---- host setup (iter_rounding_init.h:195-251) ----
        MK_CHECK(g_rows >= 0 && g_cols > 0 && g_ptr && g_ptr[0] == 0, MK_ERR_ARG, "G2");
        int64_t g_bad = 0;
#pragma omp parallel for reduction(+ : g_bad) schedule(static) if (g_rows > 100000)
        for (int i = 0; i < g_rows; i++) g_bad += g_ptr[i + 1] < g_ptr[i] ? 1 : 0;
        MK_CHECK(g_bad == 0, MK_ERR_ARG, "G2 row pointers");
        const CsrView G2{g_rows, g_cols, g_ptr, g_col, g_val};  // the caller's arrays, in place
        SprsD Cf(c_rows, c_cols);
        if (c_rows > 0) {
            Cf.ptr.assign(c_ptr, c_ptr + c_rows + 1);
            Cf.col.assign(c_col, c_col + c_ptr[c_rows]);
            Cf.val.assign(c_val, c_val + c_ptr[c_rows]);
            // explicit zeros stay: the reference collects the participating columns from every
            // stored entry of Cfull (iter_rounding_init.h:98-110), zeros of its products included
        }
        std::vector<int> uniqueJ;
        for (int64_t k = 0; k < Cf.nonZeros(); k++) uniqueJ.push_back(Cf.col[k]);
        std::sort(uniqueJ.begin(), uniqueJ.end());
        uniqueJ.erase(std::unique(uniqueJ.begin(), uniqueJ.end()), uniqueJ.end());
        std::vector<int> jmask(size_t(nfull), -1);
        for (size_t i = 0; i < uniqueJ.size(); i++) jmask[uniqueJ[i]] = int(i);
        std::vector<int> nonPart;
        metriko::MatXd US;
        int crank = 0;
        if (c_rows != 0) {
            for (int i = 0; i < nfull; i++) if (jmask[i] < 0) nonPart.push_back(i);
            metriko::MatXd CS = metriko::MatXd::Zero(c_rows, int(uniqueJ.size()));
            for (int i = 0; i < c_rows; i++)
                for (int64_t k = Cf.ptr[i]; k < Cf.ptr[i + 1]; k++) CS(i, jmask[Cf.col[k]]) = Cf.val[k];
            MK_LOG("[igm] constraints %d x %zu (of %d variables), FullPivLU...\n", c_rows, uniqueJ.size(), nfull);
            US = metriko::fullpiv_lu_kernel(CS, &crank);
            MK_LOG("[igm] kernel basis %d x %d, rank %d, %.0f ms\n", int(US.rows()), int(US.cols()), crank,
                   now_ms() - t_all);
        } else {
            nonPart.resize(size_t(nfull));
            std::iota(nonPart.begin(), nonPart.end(), 0);
            uniqueJ.clear();
        }
        const int npn = int(nonPart.size());
        const int dimker = US.rows() ? US.cols() : 0;
        const int k = npn + dimker;
        // UFull = permMat * blockdiag(I, USmall): row nonPart[i] -> e_i, row uniqueJ[r] -> USmall row r
        SprsD UF(nfull, k);
        {
            // written row by row: the nonPart rows (jmask < 0, ascending) take columns 0, 1, ...;
            // row uniqueJ[r] takes USmall row r's nonzeros at columns npn + j, ascending
            int64_t nz = 0;
            for (int i = 0; i < nfull; i++) {
                if (jmask[i] < 0) nz++;
                else for (int j = 0; j < dimker; j++) nz += US(jmask[i], j) != 0.0 ? 1 : 0;
                UF.ptr[size_t(i) + 1] = nz;
            }
            UF.col.resize(size_t(nz));
            UF.val.resize(size_t(nz));
            int np = 0;
            for (int i = 0; i < nfull; i++) {
                int64_t w = UF.ptr[size_t(i)];
                if (jmask[i] < 0) { UF.col[size_t(w)] = np++; UF.val[size_t(w)] = 1.0; continue; }
                for (int j = 0; j < dimker; j++)
                    if (US(jmask[i], j) != 0.0) { UF.col[size_t(w)] = npn + j; UF.val[size_t(w)] = US(jmask[i], j); w++; }
            }
        }
        MK_LOG("[igm] UFull %d x %d nnz %lld, %.0f ms\n", nfull, k, (long long)UF.nonZeros(), now_ms() - t_all);
        SprsD X2F = x2f_product(G2, UF, npn);  // G2 * UFull (naive, without length)
        MK_LOG("[igm] X2F nnz %lld, %.0f ms\n", (long long)X2F.nonZeros(), now_ms() - t_all);
        SprsD X2Ft = X2F.transpose();
        MK_LOG("[igm] X2F^T %.0f ms\n", now_ms() - t_all);
        SprsD E = X2Ft * X2F;
        MK_LOG("[igm] E nnz %lld, %.0f ms\n", (long long)E.nonZeros(), now_ms() - t_all);
#pragma omp parallel for schedule(static) if (E.val.size() > 200000)
        for (int64_t q = 0; q < int64_t(E.val.size()); q++) E.val[size_t(q)] *= length;   // E = X2F^T X2F * length
        std::vector<double> fr(static_cast<size_t>(k));
        X2Ft.multiply(F2, fr.data());        // f = X2F^T F2
        std::vector<int> fixIdx(fixedIdcs, fixedIdcs + nfixed);
        SprsD Cm = UF.rows_of(fixIdx);       // constMat = UFull rows at fixedIdcs
        S.k = k;
        S.kernel_rank = crank;
        S.setup_ms = now_ms() - t_all;
        MK_LOG("[igm] host setup %.0f ms (k=%d, kernel rank %d)\n", S.setup_ms, k, crank);

        // ---- Poisson KKT [[E, C^T],[C, 0]] [x; l] = [f; g] (iter_rounding_init.h:270-307) ----
        // f lies in range(E) and C only pins the translations E annihilates, so the
        // multiplier vanishes and x solves the SPD system (E + w C^T C) x = f + w C^T g.
        double t0 = now_ms();
        MK_LOG("[igm] KKT: start\n");
        double dmean = 0.0;
        {
            // diagonal entries found in parallel, summed in row order
            std::vector<double> dg(size_t(k), 0.0);
#pragma omp parallel for schedule(static, 4096) if (k > 100000)
            for (int i = 0; i < k; i++)
                for (int64_t q = E.ptr[i]; q < E.ptr[i + 1]; q++) if (E.col[q] == i) dg[size_t(i)] += E.val[q];
            for (int i = 0; i < k; i++) dmean += dg[size_t(i)];
        }
        dmean = k ? dmean / k : 1.0;
        if (!(dmean > 0.0)) dmean = 1.0;
        // Eaug = E + dmean Cm^T Cm: the penalty's few entries added to a copy of E's values
        // when they lie in E's pattern (the fixed DOFs' own entries), else the sparse sum
        const SprsD Pen = (Cm.transpose() * Cm).scaled(dmean);
        metriko::HVec<double> eaug_val;
        SprsD EaugSum;
        bool in_pattern = true;
        {
            std::vector<std::pair<int64_t, double>> adds;
            for (int i = 0; i < Pen.rows() && in_pattern; i++)
                for (int64_t q = Pen.ptr[i]; q < Pen.ptr[i + 1] && in_pattern; q++) {
                    const int* b = E.col.data() + E.ptr[i];
                    const int* e = E.col.data() + E.ptr[i + 1];
                    const int* f = std::lower_bound(b, e, Pen.col[q]);
                    if (f == e || *f != Pen.col[q]) in_pattern = false;
                    else adds.emplace_back(int64_t(f - E.col.data()), Pen.val[q]);
                }
            if (in_pattern) {
                eaug_val.resize(E.val.size());
#pragma omp parallel for schedule(static) if (E.val.size() > 200000)
                for (int64_t q = 0; q < int64_t(E.val.size()); q++) eaug_val[size_t(q)] = E.val[size_t(q)];
                for (const auto& a : adds) eaug_val[size_t(a.first)] = 1.0 * eaug_val[size_t(a.first)] + 1.0 * a.second;
            } else {
                EaugSum = E + Pen;
            }
        }
        std::vector<double> rhs(fr);
        {
            std::vector<double> ctg(size_t(k), 0.0);
            for (int i = 0; i < Cm.rows(); i++)
                for (int64_t q = Cm.ptr[i]; q < Cm.ptr[i + 1]; q++) ctg[Cm.col[q]] += Cm.val[q] * fixedVals[i];
            for (int i = 0; i < k; i++) rhs[i] += dmean * ctg[i];
        }
        OpHolder hE, hA0, hX, hXt;
        // the rounding operator A0 = wI^2 X2F^T X2F + wC^2 I, X2F = length * G2 * UFull, i.e.
        // length * E + wC^2 I (wI = 1, wC = 0.01, iter_rounding_loop.h:109); its multigrid
        // hierarchy preconditions both the KKT solve and every rounding solve.
        const double wI = 1.0, wK = 10e4, wC = 0.01;
        // device order of the unknowns (empty: the caller's): vpos[i] = device index of unknown i
        std::vector<int> vpos, vinv;
        const bool ordered = o.precond == MK_PRECOND_AMG && igm_order_wanted(k, flags);
        {
            const double ts = now_ms();
            if (ordered) {
                // the hierarchy first (on the caller's numbering), then the operator in its order
                const metriko::HVec<double> a0v = scaled_shifted(E, length * wI * wI, wC * wC);
                std::vector<double> nv = near_null(G2, CsrView::of(Cf), X2F, n, integerIdcs, nint, nonPart, uniqueJ, US);
                const int nnull = nv.empty() ? 0 : n;
                S.null_vectors = nnull;
                MK_LOG("[igm] near-null vectors %d, %.0f ms\n", nnull, now_ms() - t_all);
                mk::Mg* mg = mk::mg_build_real(c, k, E.ptr.data(), E.col.data(), a0v.data(), o, nnull ? &nv : nullptr,
                                               nnull, &vpos);
                if (!vpos.empty()) {
                    vinv.resize(vpos.size());
                    for (size_t i = 0; i < vpos.size(); i++) vinv[size_t(vpos[i])] = int(i);
                }
                try {
                    hA0.op = make_op_arrays(c, k, k, E.ptr.data(), E.col.data(), a0v.data(), o, 1.0, &vinv, &vpos);
                } catch (...) {
                    mk::mg_free(mg);
                    throw;
                }
                hA0.op->mg = mg;
                hA0.op->label = "igm_A0";
                S.device_order = vpos.empty() ? 0 : 1;
                mk::mg_set_labels(mg, hA0.op->label);
                MK_LOG("[igm] A0 multigrid and operator (device order) %.0f ms\n", now_ms() - t_all);
            } else {
                hA0.op = mk::op_create_owned(c, k, metriko::HVec<int64_t>(E.ptr), metriko::HVec<int>(E.col),
                                             scaled_shifted(E, length * wI * wI, wC * wC), o);
                MK_LOG("[igm] A0 upload %.0f ms\n", now_ms() - t_all);
                hA0.op->label = "igm_A0";
                if (o.precond == MK_PRECOND_AMG) {
                    hA0.op->nullvec = near_null(G2, CsrView::of(Cf), X2F, n, integerIdcs, nint, nonPart, uniqueJ, US);
                    hA0.op->nnull = hA0.op->nullvec.empty() ? 0 : n;
                    S.null_vectors = hA0.op->nnull;
                }
                MK_LOG("[igm] near-null vectors %d, %.0f ms\n", hA0.op->nnull, now_ms() - t_all);
                mk::ensure_mg(c, hA0.op, o);
                MK_LOG("[igm] A0 multigrid %.0f ms\n", now_ms() - t_all);
            }
            S.setup_ms += now_ms() - ts;
            t0 += now_ms() - ts;  // the KKT time excludes the A0 operator and hierarchy set-up
        }
        MK_LOG("[igm] KKT: Eaug assembled %.0f ms\n", now_ms() - t0);
        hE.op = in_pattern ? make_op_arrays(c, k, k, E.ptr.data(), E.col.data(), eaug_val.data(), o, 1.0, &vinv, &vpos)
                           : make_rect_op(c, EaugSum, o, 1.0, &vinv, &vpos);
        MK_LOG("[igm] KKT: Eaug uploaded %.0f ms\n", now_ms() - t0);
        hE.op->label = "igm_Eaug";
        DevBuf<double> dXS(static_cast<size_t>(k)), drhs;
        {
            const std::vector<double> rd = to_dev_order(vpos, rhs.data(), rhs.size());
            drhs.upload(rd.data(), rd.size(), K.st());
            K.sync();
        }
        MK_HIP(hipMemsetAsync(dXS.p, 0, size_t(k) * sizeof(double), K.st()));
        {
            const mk::SolveResult kr = K.solve(hE.op, drhs.p, dXS.p, hA0.op->mg);
            MK_LOG("[igm] KKT PCG %d iterations, residual %.2e, %.0f ms\n", kr.iters, kr.rel_residual, now_ms() - t0);
        }
        std::vector<double> XSmall(static_cast<size_t>(k));
        {
            std::vector<double> xd(static_cast<size_t>(k));
            dXS.download(xd.data(), xd.size(), K.st());
            K.sync();
            from_dev_order(vpos, xd, XSmall.data());
        }
        {   // KKT residual (host, FP64): ||E x - f|| + ||C x - g|| relative to ||f||
            std::vector<double> Ex(static_cast<size_t>(k));
            E.multiply(XSmall.data(), Ex.data());
            double rn = 0, fn = 0;
            for (int i = 0; i < k; i++) { rn += (Ex[i] - fr[i]) * (Ex[i] - fr[i]); fn += fr[i] * fr[i]; }
            for (int i = 0; i < Cm.rows(); i++) {
                double s = -fixedVals[i];
                for (int64_t q = Cm.ptr[i]; q < Cm.ptr[i + 1]; q++) s += Cm.val[q] * XSmall[Cm.col[q]];
                rn += s * s;
            }
            S.kkt_rel_residual = fn > 0 ? std::sqrt(rn / fn) : std::sqrt(rn);
        }
        S.kkt_ms = now_ms() - t0;
        MK_LOG("[igm] KKT solve %.0f ms\n", S.kkt_ms);
        std::vector<double> x0(static_cast<size_t>(nfull));
        UF.multiply(XSmall.data(), x0.data());

        *success = 1;
        int nd = 0;
        auto record = [&](int idx, int val, double margin) {
            if (nd < max_dec) {
                if (dec_index) dec_index[nd] = idx;
                if (dec_value) dec_value[nd] = val;
                if (dec_margin) dec_margin[nd] = margin;
            }
            nd++;
        };
        if (flags->local_injectivity) {
            // damped barrier LM passes (iter_rounding.h:54-113 with localInjectivity = true)
            OpHolder hB, hBt;
            hB.op = make_rect_op(c, X2F, o, length);
            hB.op->label = "igm_inj_B";
            hBt.op = make_rect_op(c, X2Ft, o, length);
            hBt.op->label = "igm_inj_Bt";
            const InjInputs in{k, nfull, g_rows, N, nF, length, &UF, &X2F, &E, &Cm, fixedVals, F2, &XSmall,
                               hA0.op, hB.op, hBt.op};
            injective_rounding(K, in, flags, singularIdcs, nsing, integerIdcs, nint, fullx, success, record, S);
        } else if (!flags->seamless) {
            std::copy(x0.begin(), x0.end(), fullx);
        } else {
            // ---- naive LM over (x, f) (iter_rounding.h:82-86), weights iter_rounding_init.h:184 ----
            t0 = now_ms();
            const double wIn = 10e3, wCn = 1.0;
            const size_t l1 = size_t(g_rows);
            // device operator length * X2F (rectangular, hybrid SELL): B x for the naive pass
            // and the field energies of the rounding passes
            OpHolder hX2F;
            hX2F.op = make_rect_op(c, X2F, o, length, nullptr, &vpos);
            hX2F.op->label = "igm_X2F";
            DevBuf<double> dF2, dBx(l1), dfstar(l1), fcur(l1);
            dF2.upload(F2, l1, K.st());
            // f* = (wI^2 B x + wC^2 F2) / (wI^2 + wC^2) at x = XSmall (the KKT solution, on the device)
            const double aB = wIn * wIn / (wIn * wIn + wCn * wCn), aF = wCn * wCn / (wIn * wIn + wCn * wCn);
            mk::spmv64(c, hX2F.op, dXS.p, dBx.p, nullptr, mk::SPMV_PLAIN, nullptr, nullptr);
            hipLaunchKernelGGL(k_lin2, dim3(nblk(l1)), dim3(256), 0, K.st(), l1, aB, dBx.p, aF, dF2.p, dfstar.p);
            // x = (XSmall, F2); the x-block of every direction is x* - x = 0, so the pass
            // only moves f: direction (0, f* - f), energies from the f-block plus the
            // constant x-terms (fixed rows of XSmall).
            double fixres = 0.0;
            for (int i = 0; i < Cm.rows(); i++) {
                double s = -fixedVals[i];
                for (int64_t q = Cm.ptr[i]; q < Cm.ptr[i + 1]; q++) s += Cm.val[q] * XSmall[Cm.col[q]];
                fixres += s * s;
            }
            DevBuf<double> pv, dd, tr;
            MK_HIP(hipMemcpyAsync(fcur.p, dF2.p, l1 * sizeof(double), hipMemcpyDeviceToDevice, K.st()));
            auto dir = [&](const double* prev, double* d) {
                hipLaunchKernelGGL(k_lin2, dim3(nblk(l1)), dim3(256), 0, K.st(), l1, 1.0, dfstar.p, -1.0, prev, d);
            };
            auto energy = [&](const double* f) {
                // wI^2 ||f - Bx||^2 + wC^2 ||f - F2||^2 + wK^2 ||C x - g||^2
                const double a = K.sqdiff(l1, 1.0, f, dBx.p);
                const double b = K.sqdiff(l1, 1.0, f, dF2.p);
                return wIn * wIn * a + wCn * wCn * b + wIn * wIn * fixres;
            };
            DevBuf<double> F2_(l1);
            // iter_cb after every naive iteration: the x-block stays at XSmall in this pass, so
            // the callback sees x0 = UFull XSmall each time
            auto post = [&](const double*) {
                return c->naive_cb ? c->naive_cb(c->naive_user, x0.data(), nfull) != 0 : false;
            };
            lm_solve(K, l1, fcur.p, F2_.p, 100, 10e-10, dir, energy, post, pv, dd, tr);
            S.naive_ms = now_ms() - t0;
            MK_LOG("[igm] naive LM %.0f ms\n", S.naive_ms);

            // ---- seamless rounding loop (iter_rounding.h:89-112, iter_rounding_loop.h) ----
            t0 = now_ms();
            // b_int = length wI^2 X2F^T F2_ with F2_ = f* (the pass ends on f*): X2F^T f* =
            // aB E XSmall + aF X2F^T F2 (E = length X2F^T X2F, X2F^T F2 = the KKT right-hand
            // side), a 9M- instead of a 30M-entry product and no download of F2_
            std::vector<double> bint(static_cast<size_t>(k));
            E.multiply(XSmall.data(), bint.data());
            for (int i = 0; i < k; i++) bint[size_t(i)] = length * wI * wI * (aB * bint[size_t(i)] + aF * fr[size_t(i)]);
            MK_LOG("[igm] rounding set-up (b_int) %.0f ms\n", now_ms() - t0);
            DevBuf<double> dfield(l1), dbint, dxcur(static_cast<size_t>(k)), dxprev(static_cast<size_t>(k)), dxt(static_cast<size_t>(k)), dxstar(static_cast<size_t>(k)),
                dxlm(static_cast<size_t>(k)), db0(static_cast<size_t>(k)), drow(static_cast<size_t>(k)), dcand, dfixv, dcoef;
            {
                const std::vector<double> bd = to_dev_order(vpos, bint.data(), bint.size());
                dbint.upload(bd.data(), bd.size(), K.st());
                K.sync();
            }
            MK_HIP(hipMemcpyAsync(dxcur.p, dXS.p, size_t(k) * sizeof(double), hipMemcpyDeviceToDevice, K.st()));
            MK_HIP(hipMemsetAsync(dxt.p, 0, size_t(k) * sizeof(double), K.st()));
            // candidate rows of UFull (all indices the loop can fix)
            std::vector<int> singular(singularIdcs, singularIdcs + nsing), integer(integerIdcs, integerIdcs + nint);
            RowSet cand;
            cand.cmap = &vpos;
            std::vector<int> candIdx;
            std::vector<int> candPos(size_t(nfull), -1);
            for (int i : singular) if (candPos[i] < 0) { candPos[i] = int(candIdx.size()); candIdx.push_back(i); cand.push(UF, i); }
            for (int i : integer) if (candPos[i] < 0) { candPos[i] = int(candIdx.size()); candIdx.push_back(i); cand.push(UF, i); }
            cand.sync(K.st());
            dcand.alloc(std::max<size_t>(candIdx.size(), 1));
            RowSet fixedRows;
            fixedRows.cmap = &vpos;
            std::vector<int> fixedI;
            std::vector<double> fixedV;
            const size_t ncand = std::max<size_t>(candIdx.size(), 1);
            const size_t nc_ = candIdx.size();
            DevBuf<double> Zall(size_t(k) * ncand);       // Woodbury column z_p = A0^{-1} u_p of candidate p
            std::vector<double> UZall(nc_ * nc_, 0.0);    // UZall[c * nc_ + p] = u_c . z_p
            std::vector<uint8_t> haveZ(nc_, 0);
            std::vector<int> left = flags->round_seams ? integer : singular;
            bool roundedSingulars = false, rounded = false;
            // Batched path: every candidate the loop can fix gets its column up front, in blocks
            // of up to 64 right-hand sides (FP64 PCG, V-cycle on the whole block: FP32, or FP64
            // in MK_PREC_FP64).
            if (o.precond == MK_PRECOND_AMG && hA0.op->mg && nc_ > 0) {
                std::vector<int> need;  // candidate positions, in candidate order
                std::vector<uint8_t> mark(nc_, 0);
                for (int i : left) mark[size_t(candPos[i])] = 1;
                if (!flags->round_seams) for (int i : integer) mark[size_t(candPos[i])] = 1;
                for (size_t p = 0; p < nc_; p++) if (mark[p]) need.push_back(int(p));
                // blocks of at most 64 columns, each as narrow as its column count allows
                size_t mwmax = need.size() > 32 ? 64 : need.size() > 16 ? 32 : need.size() > 8 ? 16 : 8;
                if (const char* e = std::getenv("METRIKO_MAX_BLOCK"))
                    mwmax = std::min<size_t>(mwmax, size_t(std::max(8, std::atoi(e))));
                DevBuf<double> Bp(size_t(k) * mwmax), Xp(size_t(k) * mwmax);
                DevBuf<int> dsel(mwmax);
                for (size_t b0 = 0, cnt_ = 0; b0 < need.size(); b0 += cnt_) {
                    cnt_ = std::min<size_t>(mwmax, need.size() - b0);
                    const int cnt = int(cnt_);
                    const int mw = cnt > 32 ? 64 : cnt > 16 ? 32 : cnt > 8 ? 16 : 8;
                    std::vector<int> sel(size_t(mw), -1);
                    for (int j = 0; j < cnt; j++) sel[size_t(j)] = need[b0 + size_t(j)];
                    dsel.upload(sel.data(), sel.size(), K.st());
                    hipLaunchKernelGGL(k_zero, dim3(nblk(size_t(k) * mw)), dim3(256), 0, K.st(), size_t(k) * mw, Bp.p);
                    hipLaunchKernelGGL(k_scatter_rows_packed, dim3(unsigned(mw)), dim3(64), 0, K.st(), mw, dsel.p,
                                       cand.ptr.p, cand.col.p, cand.val.p, Bp.p);
                    mk::BlockResult br = mk::block_pcg(c, hA0.op, hA0.op->mg, mw, Bp.p, Xp.p, o.tol, o.maxiter,
                                                       o.check_every, o.precision == MK_PREC_FP64);
                    K.acc.iterations += br.iters;
                    K.acc.solves += cnt;
                    S.batched_solves += cnt;
                    S.batch_iterations += br.iters;
                    if (br.status != MK_OK)
                        throw Error(br.status, "IGM batched solve failed: " + std::string(mk_status_str(br.status)));
                    hipLaunchKernelGGL(k_unpack_cols, dim3(nblk(size_t(k) * mw)), dim3(256), 0, K.st(), size_t(k), mw,
                                       dsel.p, (const double*)Xp.p, Zall.p);
                    for (int j = 0; j < cnt; j++) haveZ[size_t(need[b0 + size_t(j)])] = 1;
                    MK_LOG("[igm] batched Woodbury columns %zu/%zu: %d iterations\n", b0 + size_t(cnt), need.size(),
                           br.iters);
                }
                // U Z for every candidate pair, on the device
                DevBuf<double> duz(nc_ * nc_);
                hipLaunchKernelGGL(k_uz, dim3(nblk(nc_ * nc_)), dim3(256), 0, K.st(), int(nc_), size_t(k), cand.ptr.p,
                                   cand.col.p, cand.val.p, (const double*)Zall.p, duz.p);
                duz.download(UZall.data(), UZall.size(), K.st());
                K.sync();
                MK_LOG("[igm] batched Woodbury columns done %.0f ms\n", now_ms() - t0);
            }
            DevBuf<int> dzidx;
            std::vector<double> Xc(candIdx.size());
            // ||F2_ - length X2F x||^2 (residual SpMV with fused norm)
            auto field_energy = [&](const double* x) {
                mk::spmv64(c, hX2F.op, x, dfield.p, F2_.p, mk::SPMV_RESID, K.part.p + 2048, nullptr);
                double h = 0.0;
                MK_HIP(hipMemcpyAsync(&h, K.part.p + 2048, sizeof(double), hipMemcpyDeviceToHost, K.st()));
                K.sync();
                return h;
            };
            std::vector<double> dotbuf;
            DevBuf<double> ddots;
            auto fixed_dots = [&](const double* x, std::vector<double>& out) {
                out.assign(size_t(fixedRows.nrows), 0.0);
                if (!fixedRows.nrows) return;
                ddots.ensure(size_t(fixedRows.nrows));
                K.rowdots(fixedRows, x, ddots.p);
                ddots.download(out.data(), out.size(), K.st());
                K.sync();
            };
            int lm_rounds = 0, rounds = 0, last_r = 0;
            DevBuf<double> dxprev_last;
            DevBuf<double> pv2, dd2, tr2;
            double tm_xt = 0.0, tm_lm = 0.0, tm_all0 = now_ms();
            int xt_iters = 0;
            while (!left.empty()) {
                // initFixedIndices (iter_rounding_loop.h:29-74)
                MK_HIP(hipMemcpyAsync(dxprev.p, dxcur.p, size_t(k) * sizeof(double), hipMemcpyDeviceToDevice, K.st()));
                K.rowdots(cand, dxcur.p, dcand.p);
                dcand.download(Xc.data(), Xc.size(), K.st());
                K.sync();
                double minDif = 3276700.0;
                int minIdx = -1;
                for (size_t i = 0; i < left.size(); i++) {
                    const double v = Xc[size_t(candPos[left[i]])];
                    const double dd = std::fabs(v - std::round(v));
                    if (dd < minDif) { minIdx = int(i); minDif = dd; }
                }
                const int pick = left[size_t(minIdx)];
                const double Xp = Xc[size_t(candPos[pick])];
                fixedI.push_back(pick);
                fixedV.push_back(double(int(std::round(Xp))));
                record(pick, int(std::round(Xp)), minDif);
                left.erase(left.begin() + minIdx);
                fixedRows.push(UF, pick);
                fixedRows.sync(K.st());
                // Woodbury column z = A0^{-1} u_pick (computed now unless the batch has it)
                const int cp = candPos[pick];
                if (!haveZ[size_t(cp)]) {
                    double* z = Zall.p + size_t(cp) * size_t(k);
                    hipLaunchKernelGGL(k_zero, dim3(nblk(static_cast<size_t>(k))), dim3(256), 0, K.st(), size_t(k), drow.p);
                    const int64_t b = cand.hptr[size_t(cp)], e = cand.hptr[size_t(cp) + 1];
                    if (e > b)
                        hipLaunchKernelGGL(k_scatter_row, dim3(nblk(size_t(e - b))), dim3(256), 0, K.st(), b, e,
                                           cand.col.p, cand.val.p, drow.p);
                    hipLaunchKernelGGL(k_zero, dim3(nblk(static_cast<size_t>(k))), dim3(256), 0, K.st(), size_t(k), z);
                    K.solve(hA0.op, drow.p, z);
                    std::vector<double> cd(nc_);
                    K.rowdots(cand, z, dcand.p);
                    dcand.download(cd.data(), cd.size(), K.st());
                    K.sync();
                    for (size_t q = 0; q < nc_; q++) UZall[q * nc_ + size_t(cp)] = cd[q];
                    haveZ[size_t(cp)] = 1;
                }
                if (left.empty() && !flags->round_seams && !roundedSingulars) {
                    left = integer;
                    roundedSingulars = true;
                }
                rounds++;
                if (rounds % 10 == 0) MK_LOG("[igm] rounding: %d rounds, %d left\n", rounds, int(left.size()));
                if (!(minDif > 10e-7)) continue;
                rounded = true;
                lm_rounds++;
                const int r = int(fixedI.size());
                if (c->keep_igm_state) {  // the last LM round's system is kept for the sized solves
                    last_r = r;
                    dxprev_last.ensure(static_cast<size_t>(k));
                    MK_HIP(hipMemcpyAsync(dxprev_last.p, dxprev.p, size_t(k) * sizeof(double), hipMemcpyDeviceToDevice,
                                          K.st()));
                }
                std::vector<double> UZ(size_t(r) * size_t(r));
                std::vector<int> zidx(static_cast<size_t>(r));
                for (int a = 0; a < r; a++) {
                    zidx[size_t(a)] = candPos[fixedI[size_t(a)]];
                    for (int bq = 0; bq < r; bq++)
                        UZ[size_t(a * r + bq)] = UZall[size_t(candPos[fixedI[size_t(a)]]) * nc_ + size_t(candPos[fixedI[size_t(bq)]])];
                }
                dzidx.upload(zidx.data(), zidx.size(), K.st());
                // xt = A0^{-1} (b_int + wC^2 xPrev)  (warm start: previous xt)
                hipLaunchKernelGGL(k_lin2, dim3(nblk(static_cast<size_t>(k))), dim3(256), 0, K.st(), size_t(k), 1.0, dbint.p,
                                   wC * wC, dxprev.p, db0.p);
                {
                    const double t1 = now_ms();
                    xt_iters += K.solve(hA0.op, db0.p, dxt.p).iters;
                    tm_xt += now_ms() - t1;
                }
                std::vector<double> Uxt;
                fixed_dots(dxt.p, Uxt);
                std::vector<double> M(size_t(r) * size_t(r)), rhs2(static_cast<size_t>(r));
                for (int i = 0; i < r; i++) {
                    for (int j = 0; j < r; j++) M[size_t(i * r + j)] = UZ[size_t(i * r + j)];
                    M[size_t(i * r + i)] += 1.0 / (wK * wK);
                    rhs2[size_t(i)] = fixedV[size_t(i)] - Uxt[size_t(i)];
                }
                std::vector<double> cc = dense_solve(M, rhs2, r);
                // x* = xt + Z c
                dcoef.upload(cc.data(), cc.size(), K.st());
                hipLaunchKernelGGL(k_combine, dim3(nblk(static_cast<size_t>(k))), dim3(256), 0, K.st(), size_t(k), r, dxt.p, Zall.p,
                                   (const int*)dzidx.p, dcoef.p, dxstar.p);
                std::vector<double> fr_(fixedV);
                auto dir = [&](const double* prev, double* d) {
                    hipLaunchKernelGGL(k_lin2, dim3(nblk(static_cast<size_t>(k))), dim3(256), 0, K.st(), size_t(k), 1.0, dxstar.p,
                                       -1.0, prev, d);
                };
                auto energy = [&](const double* x) {
                    const double a = field_energy(x);
                    const double b = K.sqdiff(size_t(k), 1.0, x, dxprev.p);
                    std::vector<double> ux;
                    fixed_dots(x, ux);
                    double q = 0.0;
                    for (int i = 0; i < r; i++) q += (ux[size_t(i)] - fr_[size_t(i)]) * (ux[size_t(i)] - fr_[size_t(i)]);
                    return wI * wI * a + wC * wC * b + wK * wK * q;
                };
                {
                    const double t1 = now_ms();
                    lm_solve(K, size_t(k), dxcur.p, dxlm.p, 100, 1e-7, dir, energy, [](const double*) { return false; },
                             pv2, dd2, tr2);
                    tm_lm += now_ms() - t1;
                }
                // post_checking (iter_rounding_loop.h:89-100)
                std::vector<double> ux;
                fixed_dots(dxlm.p, ux);
                double mx = -1.0;
                for (int i = 0; i < r; i++) mx = std::max(mx, std::abs(ux[size_t(i)] - std::round(ux[size_t(i)])));
                MK_HIP(hipMemcpyAsync(dxcur.p, dxlm.p, size_t(k) * sizeof(double), hipMemcpyDeviceToDevice, K.st()));
                if (!(mx <= 10e-7)) { *success = 0; break; }
            }
            MK_LOG("[igm] rounding loop %.0f ms: %d rounds, %d with LM; xt solves %.0f ms (%d iterations), LM "
                   "passes %.0f ms\n", now_ms() - tm_all0, rounds, lm_rounds, tm_xt, xt_iters, tm_lm);
            S.rounds = rounds;
            S.lm_rounds = lm_rounds;
            S.n_fixed = int(fixedI.size());
            if (rounded) {
                std::vector<double> xd(static_cast<size_t>(k)), xs(static_cast<size_t>(k));
                dxcur.download(xd.data(), xd.size(), K.st());
                K.sync();
                from_dev_order(vpos, xd, xs.data());
                UF.multiply(xs.data(), fullx);
            } else {
                std::copy(x0.begin(), x0.end(), fullx);
            }
            S.rounding_ms = now_ms() - t0;
            if (c->keep_igm_state && rounded && last_r > 0) {
                // the last LM round's system (A0 + wK^2 U^T U) x = length wI^2 X2F^T F2 + wC^2 xPrev
                // + wK^2 U^T r, with its Woodbury columns, for mk_igm_sized_uv / _grad
                auto* st = new mk_igm_state();
                st->ctx = c;
                st->o = o;
                st->k = k;
                st->nfull = nfull;
                st->l1 = int(l1);
                st->nF = nF;
                st->N = N;
                st->r = last_r;
                st->length = length;
                st->wI = wI;
                st->wK = wK;
                st->wC = wC;
                st->A0 = hA0.op;
                hA0.op = nullptr;
                st->X2F = hX2F.op;
                hX2F.op = nullptr;
                st->UF = std::move(UF);
                st->X2Ft = std::move(X2Ft);
                st->pos = vpos;
                st->F2.assign(F2, F2 + l1);
                st->fixedV.assign(fixedV.begin(), fixedV.begin() + last_r);
                st->UZ.resize(size_t(last_r) * size_t(last_r));
                st->Urows.resize(size_t(last_r));
                st->Z.alloc(size_t(k) * size_t(last_r));
                for (int a = 0; a < last_r; a++) {
                    const int pa = candPos[fixedI[size_t(a)]];
                    st->Urows[size_t(a)] = fixedI[size_t(a)];
                    for (int bq = 0; bq < last_r; bq++)
                        st->UZ[size_t(a) * last_r + bq] = UZall[size_t(pa) * nc_ + size_t(candPos[fixedI[size_t(bq)]])];
                    MK_HIP(hipMemcpyAsync(st->Z.p + size_t(a) * size_t(k), Zall.p + size_t(pa) * size_t(k),
                                          size_t(k) * sizeof(double), hipMemcpyDeviceToDevice, K.st()));
                }
                st->xprev.alloc(static_cast<size_t>(k));
                MK_HIP(hipMemcpyAsync(st->xprev.p, dxprev_last.p, size_t(k) * sizeof(double), hipMemcpyDeviceToDevice,
                                      K.st()));
                K.sync();
                if (c->igm_state) mk_igm_state_destroy(c->igm_state);
                c->igm_state = st;
            }
        }
        if (ndec) *ndec = nd;
        K.sync();
        S.solve = K.acc;
        mk::stats_fill(c, &S.solve);
        S.solve.ms = now_ms() - t_all;
        if (stats) *stats = S;
        return MK_OK;
    } catch (const Error& e) {
        c->last_error = e.what();
        return e.code;
    } catch (const std::exception& e) {
        c->last_error = e.what();
        return MK_ERR_ARG;
    }
}

namespace {
// U y for the kept fixed rows (device), to the host
std::vector<double> state_rowdots(mk_igm_state* st, const double* y) {
    mk_ctx* c = st->ctx;
    if (!st->uptr.p) {
        std::vector<int64_t> hp{0};
        std::vector<int> hc;
        std::vector<double> hv;
        for (int row : st->Urows) {
            for (int64_t q = st->UF.ptr[row]; q < st->UF.ptr[row + 1]; q++) {
                hc.push_back(st->pos.empty() ? st->UF.col[q] : st->pos[size_t(st->UF.col[q])]);
                hv.push_back(st->UF.val[q]);
            }
            hp.push_back(int64_t(hc.size()));
        }
        st->uptr.upload(hp.data(), hp.size(), c->stream);
        st->ucol.upload(hc.data(), std::max<size_t>(hc.size(), 1), c->stream);
        st->uval.upload(hv.data(), std::max<size_t>(hv.size(), 1), c->stream);
    }
    DevBuf<double> out(static_cast<size_t>(st->r));
    hipLaunchKernelGGL(k_rowset_dot, dim3(st->r), dim3(256), 0, c->stream, st->r, (const int64_t*)st->uptr.p,
                       (const int*)st->ucol.p, (const double*)st->uval.p, y, out.p);
    std::vector<double> h(static_cast<size_t>(st->r));
    out.download(h.data(), h.size(), c->stream);
    MK_HIP(hipStreamSynchronize(c->stream));
    return h;
}
// M^{-1} v with M = U Z + I / wK^2 (r x r)
std::vector<double> state_msolve(const mk_igm_state* st, const std::vector<double>& v) {
    const int r = st->r;
    std::vector<double> M(st->UZ);
    for (int i = 0; i < r; i++) M[size_t(i) * r + i] += 1.0 / (st->wK * st->wK);
    return dense_solve(M, v, r);
}
// x = y + Z c (device)
void state_combine(mk_igm_state* st, const double* y, const std::vector<double>& cf, double* x) {
    mk_ctx* c = st->ctx;
    std::vector<int> idx(static_cast<size_t>(st->r));
    std::iota(idx.begin(), idx.end(), 0);
    DevBuf<int> didx;
    DevBuf<double> dc;
    didx.upload(idx.data(), idx.size(), c->stream);
    dc.upload(cf.data(), cf.size(), c->stream);
    hipLaunchKernelGGL(k_combine, dim3(nblk(size_t(st->k))), dim3(256), 0, c->stream, size_t(st->k), st->r, y,
                       (const double*)st->Z.p, (const int*)didx.p, (const double*)dc.p, x);
    MK_HIP(hipStreamSynchronize(c->stream));
}
void check_scale(const mk_igm_state* st, const double* face_scale) {
    for (int f = 0; f < st->nF; f++)
        MK_CHECK(face_scale[f] == face_scale[f] && face_scale[f] > 0.0, MK_ERR_ARG, "face_scale must be positive");
}
template <typename Fn>
mk_status state_guard(mk_ctx* c, Fn&& fn) {
    try {
        fn();
        return MK_OK;
    } catch (const Error& e) {
        if (c) c->last_error = e.what();
        return e.code;
    } catch (const std::exception& e) {
        if (c) c->last_error = e.what();
        return MK_ERR_ARG;
    }
}
}  // namespace

extern "C" {

mk_status mk_igm_injectivity_check(mk_ctx* c, int N, int nF, const int32_t* cut_idx, int f_rows, int f_cols,
                                   const int64_t* f_ptr, const int32_t* f_col, const double* f_val) {
    if (!c) return MK_ERR_ARG;
    if (nF <= 0) {
        c->inj_check = mk_inj_check_data{};
        return MK_OK;
    }
    if (N < 2 || !cut_idx || f_rows <= 0 || f_cols <= 0 || f_rows % N || !f_ptr || f_ptr[0] != 0) return MK_ERR_ARG;
    const int nV = f_rows / N;
    for (int64_t i = 0; i < 3 * int64_t(nF); i++)
        if (cut_idx[i] < 0 || cut_idx[i] >= nV) { c->last_error = "cut_idx out of range"; return MK_ERR_ARG; }
    for (int i = 0; i < f_rows; i++) if (f_ptr[i + 1] < f_ptr[i]) return MK_ERR_ARG;
    for (int64_t q = 0; q < f_ptr[f_rows]; q++)
        if (f_col[q] < 0 || f_col[q] >= f_cols) { c->last_error = "fullx2Nfunc column out of range"; return MK_ERR_ARG; }
    mk_inj_check_data d;
    d.N = N;
    d.nF = nF;
    d.rows = f_rows;
    d.cols = f_cols;
    d.idx.assign(cut_idx, cut_idx + 3 * int64_t(nF));
    d.ptr.assign(f_ptr, f_ptr + f_rows + 1);
    d.col.assign(f_col, f_col + f_ptr[f_rows]);
    d.val.assign(f_val, f_val + f_ptr[f_rows]);
    c->inj_check = std::move(d);
    return MK_OK;
}

mk_status mk_set_naive_iter_callback(mk_ctx* c, mk_naive_iter_cb cb, void* user) {
    if (!c) return MK_ERR_ARG;
    c->naive_cb = cb;
    c->naive_user = cb ? user : nullptr;
    return MK_OK;
}

mk_status mk_keep_igm_state(mk_ctx* c, int on) {
    if (!c) return MK_ERR_ARG;
    c->keep_igm_state = on != 0;
    return MK_OK;
}

mk_status mk_igm_state_take(mk_ctx* c, mk_igm_state** st) {
    if (!c || !st) return MK_ERR_ARG;
    *st = c->igm_state;
    c->igm_state = nullptr;
    return *st ? MK_OK : MK_ERR_ARG;
}

mk_status mk_igm_state_destroy(mk_igm_state* st) {
    if (!st) return MK_ERR_ARG;
    (void)hipSetDevice(st->ctx->device);
    delete st;
    return MK_OK;
}

mk_status mk_igm_state_sizes(const mk_igm_state* st, int* nF, int* nfull, int* k, int* nfixed, int* has_corner_map) {
    if (!st) return MK_ERR_ARG;
    if (nF) *nF = st->nF;
    if (nfull) *nfull = st->nfull;
    if (k) *k = st->k;
    if (nfixed) *nfixed = st->r;
    if (has_corner_map) *has_corner_map = st->has_P ? 1 : 0;
    return MK_OK;
}

mk_status mk_igm_sized_uv(mk_ctx* c, mk_igm_state* st, const double* face_scale, double* fullx, double* uv) {
    if (!c || !st || st->ctx != c || !face_scale) return MK_ERR_ARG;
    return state_guard(c, [&] {
        MK_HIP(hipSetDevice(c->device));
        check_scale(st, face_scale);
        MK_CHECK(!uv || st->has_P, MK_ERR_ARG, "no corner map: the state did not come from mk_pipeline_run");
        const int k = st->k;
        // b = length wI^2 X2F^T (sigma .* F2) + wC^2 xPrev
        std::vector<double> sf(static_cast<size_t>(st->l1)), b(static_cast<size_t>(k));
        const int per = 2 * st->N;
        for (int f = 0; f < st->nF; f++)
            for (int j = 0; j < per; j++) sf[size_t(f) * per + j] = face_scale[f] * st->F2[size_t(f) * per + j];
        st->X2Ft.multiply(sf.data(), b.data());
        for (auto& v : b) v *= st->length * st->wI * st->wI;
        DevBuf<double> db, dxt(static_cast<size_t>(k)), dx(static_cast<size_t>(k));
        const std::vector<double> bd = to_dev_order(st->pos, b.data(), b.size());
        db.upload(bd.data(), bd.size(), c->stream);
        hipLaunchKernelGGL(k_lin2, dim3(nblk(size_t(k))), dim3(256), 0, c->stream, size_t(k), 1.0, (const double*)db.p,
                           st->wC * st->wC, (const double*)st->xprev.p, db.p);
        MK_HIP(hipMemcpyAsync(dxt.p, st->xprev.p, size_t(k) * sizeof(double), hipMemcpyDeviceToDevice, c->stream));
        const mk::SolveResult sr = mk::solve_spd(c, st->A0, db.p, dxt.p, st->o);
        MK_CHECK(sr.status == MK_OK, sr.status, "sized solve: " + std::string(mk_status_str(sr.status)));
        // x = xt + Z M^{-1} (r - U xt)
        const std::vector<double> ux = state_rowdots(st, dxt.p);
        std::vector<double> rhs(static_cast<size_t>(st->r));
        for (int i = 0; i < st->r; i++) rhs[size_t(i)] = st->fixedV[size_t(i)] - ux[size_t(i)];
        state_combine(st, dxt.p, state_msolve(st, rhs), dx.p);
        std::vector<double> xd(static_cast<size_t>(k)), xs(static_cast<size_t>(k)), fx(static_cast<size_t>(st->nfull));
        dx.download(xd.data(), xd.size(), c->stream);
        MK_HIP(hipStreamSynchronize(c->stream));
        from_dev_order(st->pos, xd, xs.data());
        st->UF.multiply(xs.data(), fx.data());
        if (fullx) std::copy(fx.begin(), fx.end(), fullx);
        if (uv) st->P.multiply(fx.data(), uv);
    });
}

mk_status mk_igm_sized_uv_grad(mk_ctx* c, mk_igm_state* st, const double* face_scale, const double* g_uv,
                               double* g_scale) {
    if (!c || !st || st->ctx != c || !face_scale || !g_uv || !g_scale) return MK_ERR_ARG;
    return state_guard(c, [&] {
        MK_HIP(hipSetDevice(c->device));
        check_scale(st, face_scale);
        MK_CHECK(st->has_P, MK_ERR_ARG, "no corner map: the state did not come from mk_pipeline_run");
        const int k = st->k;
        // g = UFull^T P^T g_uv (host, transposed products)
        std::vector<double> gf(static_cast<size_t>(st->nfull), 0.0), g(static_cast<size_t>(k), 0.0);
        for (int i = 0; i < st->P.rows(); i++)
            for (int64_t q = st->P.ptr[i]; q < st->P.ptr[i + 1]; q++) gf[size_t(st->P.col[q])] += st->P.val[q] * g_uv[i];
        for (int i = 0; i < st->UF.rows(); i++)
            for (int64_t q = st->UF.ptr[i]; q < st->UF.ptr[i + 1]; q++) g[size_t(st->UF.col[q])] += st->UF.val[q] * gf[size_t(i)];
        // adjoint: lambda = H^{-1} g = A0^{-1} g - Z M^{-1} U A0^{-1} g
        DevBuf<double> dg, dy(static_cast<size_t>(k)), dl(static_cast<size_t>(k)), dw(static_cast<size_t>(st->l1));
        const std::vector<double> gd = to_dev_order(st->pos, g.data(), g.size());
        dg.upload(gd.data(), gd.size(), c->stream);
        MK_HIP(hipMemsetAsync(dy.p, 0, size_t(k) * sizeof(double), c->stream));
        const mk::SolveResult sr = mk::solve_spd(c, st->A0, dg.p, dy.p, st->o);
        MK_CHECK(sr.status == MK_OK, sr.status, "adjoint solve: " + std::string(mk_status_str(sr.status)));
        std::vector<double> d = state_msolve(st, state_rowdots(st, dy.p));
        for (auto& v : d) v = -v;
        state_combine(st, dy.p, d, dl.p);
        // d/d sigma_f of lambda^T b = wI^2 (length X2F lambda)_(rows of f) . F2_(rows of f)
        mk::spmv64(c, st->X2F, dl.p, dw.p, nullptr, mk::SPMV_PLAIN, nullptr, nullptr);
        std::vector<double> w(static_cast<size_t>(st->l1));
        dw.download(w.data(), w.size(), c->stream);
        MK_HIP(hipStreamSynchronize(c->stream));
        const int per = 2 * st->N;
        for (int f = 0; f < st->nF; f++) {
            double a = 0.0;
            for (int j = 0; j < per; j++) a += w[size_t(f) * per + j] * st->F2[size_t(f) * per + j];
            g_scale[f] = st->wI * st->wI * a;
        }
    });
}

}  // extern "C"
