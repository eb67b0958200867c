// Smoothed-aggregation algebraic multigrid, used as the PCG preconditioner of the
// field Laplacian (complex Hermitian) and of the IGM Poisson operators (real SPD).
//
// Setup (host, FP64): strength of connection |a_ij|^2 >= theta^2 |a_ii||a_jj|; greedy
// aggregation (seed / attach / leftover passes); tentative prolongator from a near-null
// vector built per aggregate by following the operator's own phases along a BFS tree
// (B_j = -conj(a_ij)/|a_ij| B_i: parallel transport for the connection Laplacian,
// constants for a plain Laplacian); Jacobi-smoothed prolongator
// P = (I - 4/(3 rho) D^-1 A) P_tent; Galerkin coarse operator P^H A P; a dense inverse on
// the coarsest level.
// Solve (device, FP64 or FP32): symmetric V-cycle with Chebyshev(2) pre/post smoothing
// in D^-1 A on [rho/30, 1.1 rho]; every operator is a SELL-C-sigma matrix and every
// kernel honours the CG "done" flag so a converged solve costs nothing further.
#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <complex>
#include <numeric>
#include <random>
#include "mk_internal.h"
#include "metriko/core/common/sparse.h"

namespace mk {

template <typename T>
void spmv_launch(mk_ctx* c, const mk_op* A, const T* val, const T* x, T* y, const T* b, SpmvMode mode, double* out,
                 double bytes, const int* skip);

struct MgLevel {
    mk_op* A = nullptr;
    mk_op* P = nullptr;  // n_l x n_{l+1}
    mk_op* R = nullptr;  // n_{l+1} x n_l
    mk_op* RR = nullptr; // n_{l+1} x 2 n_l: [R | -R A] (restriction of the residual in one product)
    double lmax = 2.0;
    int n = 0;
    DevBuf<double> b64, x64, r64, d64, y64;  // y: second iterate buffer of the fused smoother
    DevBuf<float> b32, x32, r32, d32, y32;
    DevBuf<float> bb32, bx32, br32, by32;  // n x mw blocks for the batched V-cycle (FP32)
    DevBuf<double> bb64, bx64, br64, by64;  // the same for the FP64 batched V-cycle
};

struct Mg {
    const uint64_t uid = next_uid();  // keys the PCG graphs captured through this hierarchy
    bool cplx = false;
    std::vector<MgLevel> lv;
    int nc = 0;
    DevBuf<double> inv64;  // dense coarsest inverse (row-major, complex interleaved)
    DevBuf<float> inv32;
    int deg = 2;
    double setup_ms = 0.0;
    int64_t nnz_total = 0;
};

void mg_free(Mg* g) {
    if (!g) return;
    for (auto& L : g->lv) {
        delete L.A;
        delete L.P;
        delete L.R;
        delete L.RR;
    }
    delete g;
}

// Multigrid parameters (METRIKO_MG_* environment overrides, read once).
struct MgTune {
    double theta_real = 0.05, theta_cplx = 0.08, lmin_div = 10.0, lmax_fac = 1.2;
    int deg = 2, power = 15, phase_real = 1, coarse_real = 500, coarse_cplx = 400, smooth_long = 1, compat = 1,
        l1 = 1, attach = 1, rr = 0;
};
static const MgTune& tune() {
    static const MgTune t = [] {
        MgTune x;
        auto f = [](const char* n, double& v) { if (const char* s = std::getenv(n)) v = std::atof(s); };
        auto i = [](const char* n, int& v) { if (const char* s = std::getenv(n)) v = std::atoi(s); };
        f("METRIKO_MG_THETA_REAL", x.theta_real);
        f("METRIKO_MG_THETA_CPLX", x.theta_cplx);
        f("METRIKO_MG_LMIN_DIV", x.lmin_div);
        f("METRIKO_MG_LMAX_FAC", x.lmax_fac);
        i("METRIKO_MG_DEG", x.deg);
        i("METRIKO_MG_POWER", x.power);
        i("METRIKO_MG_PHASE_REAL", x.phase_real);
        i("METRIKO_MG_COARSE_REAL", x.coarse_real);
        i("METRIKO_MG_COARSE_CPLX", x.coarse_cplx);
        i("METRIKO_MG_SMOOTH_LONG", x.smooth_long);
        i("METRIKO_MG_COMPAT", x.compat);
        i("METRIKO_MG_L1", x.l1);
        i("METRIKO_MG_ATTACH", x.attach);
        i("METRIKO_MG_RR", x.rr);
        return x;
    }();
    return t;
}

namespace {

using metriko::Csr;
using cd = std::complex<double>;

inline double mag2(double a) { return a * a; }
inline double mag2(const cd& a) { return std::norm(a); }
inline double mag(double a) { return std::abs(a); }
inline double mag(const cd& a) { return std::abs(a); }
inline double conjv(double a) { return a; }
inline cd conjv(const cd& a) { return std::conj(a); }
inline double re(double a) { return a; }
inline double re(const cd& a) { return a.real(); }

template <typename T>
Csr<T> conj_transpose(const Csr<T>& A) {
    Csr<T> t = A.transpose();
    if constexpr (!std::is_same<T, double>::value) {
        const int64_t nz = t.nonZeros();
#pragma omp parallel for schedule(static) if (nz > 200000)
        for (int64_t k = 0; k < nz; k++) t.val[size_t(k)] = conjv(t.val[size_t(k)]);
    }
    return t;
}

template <typename T>
std::vector<double> diag_of(const Csr<T>& A) {
    std::vector<double> d(size_t(A.nr), 0.0);
#pragma omp parallel for schedule(static, 4096) if (A.nr > 100000)
    for (int i = 0; i < A.nr; i++)
        for (int64_t k = A.ptr[i]; k < A.ptr[i + 1]; k++) if (A.col[k] == i) d[i] = re(A.val[k]);
    return d;
}

// sum of f(i) over [0, n) in 64 fixed chunks (parallel, same result for any thread count)
template <typename F>
double chunked_sum(int n, F&& f) {
    constexpr int kChunks = 64;
    double part[kChunks];
    const int step = (n + kChunks - 1) / kChunks;
#pragma omp parallel for schedule(static, 1) if (n > 100000)
    for (int c = 0; c < kChunks; c++) {
        double s = 0.0;
        const int e = std::min(n, (c + 1) * step);
        for (int i = c * step; i < e; i++) s += f(i);
        part[c] = s;
    }
    double t = 0.0;
    for (int c = 0; c < kChunks; c++) t += part[c];
    return t;
}

// spectral radius of D^-1 A by power iteration (fixed seed)
template <typename T>
double rho_dinv_a(const Csr<T>& A, const std::vector<double>& d) {
    const int n = A.nr;
    std::mt19937_64 g(12345);
    // uniform on [-1, 1): the engine's 64 bits scaled by 2^-64 (values that round to 1 take the
    // largest double below it), times 2, minus 1 -- the value uniform_real_distribution<double>
    // gives for a 64-bit engine, without its per-draw long-double set-up
    auto U = [&g]() {
        double u = double(g()) * 0x1p-64;
        if (u >= 1.0) u = std::nextafter(1.0, 0.0);
        return u * 2.0 + -1.0;
    };
    metriko::HVec<T> x(static_cast<size_t>(n)), y(static_cast<size_t>(n));
    for (auto& v : x) {
        if constexpr (std::is_same<T, double>::value) {
            v = U();
        } else {
            const double re = U();
            const double im = U();
            v = cd(re, im);
        }
    }
    double lam = 1.0;
    for (int it = 0; it < tune().power; it++) {
        A.multiply(x.data(), y.data());
        double nrm = chunked_sum(n, [&](int i) {
            y[i] = y[i] / (d[i] != 0.0 ? d[i] : 1.0);
            return mag2(y[i]);
        });
        nrm = std::sqrt(nrm);
        double xn = std::sqrt(chunked_sum(n, [&](int i) { return mag2(x[i]); }));
        lam = nrm / (xn > 0 ? xn : 1.0);
        if (nrm == 0.0) break;
#pragma omp parallel for schedule(static, 4096) if (n > 100000)
        for (int i = 0; i < n; i++) x[i] = y[i] / nrm;
    }
    return lam;
}

// A <- (A + A^H) / 2, each entry computed as 0.5 a_ij + 0.5 conj(a_ji).  In place when the
// pattern is structurally symmetric (R A P of a structurally symmetric A with R = P^H always
// is): the owner of each pair i < j finds (j, i) by binary search in row j (sorted columns) and
// writes both entries, so no thread touches another's pair.  Otherwise through the transpose.
template <typename T>
void hermitian_average(Csr<T>& A) {
    const int n = A.nr;
    auto partner = [&](int i, int j) -> int64_t {  // position of (j, i), or -1
        const int* b = A.col.data() + A.ptr[size_t(j)];
        const int* e = A.col.data() + A.ptr[size_t(j) + 1];
        const int* p = std::lower_bound(b, e, i);
        return (p != e && *p == i) ? int64_t(p - A.col.data()) : -1;
    };
    std::atomic<int> asym{0};
#pragma omp parallel for schedule(static, 4096) if (A.nonZeros() > 200000)
    for (int i = 0; i < n; i++)
        for (int64_t k = A.ptr[size_t(i)]; k < A.ptr[size_t(i) + 1]; k++)
            if (A.col[size_t(k)] > i && partner(i, A.col[size_t(k)]) < 0) asym.store(1, std::memory_order_relaxed);
    if (asym.load() || A.nr != A.nc) {
        Csr<T> AH = conj_transpose(A);
        A = metriko::add(A, T(0.5), AH, T(0.5));
        return;
    }
#pragma omp parallel for schedule(static, 4096) if (A.nonZeros() > 200000)
    for (int i = 0; i < n; i++)
        for (int64_t k = A.ptr[size_t(i)]; k < A.ptr[size_t(i) + 1]; k++) {
            const int j = A.col[size_t(k)];
            if (j < i) continue;
            if (j == i) {
                const T a = A.val[size_t(k)];
                A.val[size_t(k)] = T(0.5) * a + T(0.5) * conjv(a);
                continue;
            }
            const int64_t q = partner(i, j);
            const T aij = A.val[size_t(k)], aji = A.val[size_t(q)];
            A.val[size_t(k)] = T(0.5) * aij + T(0.5) * conjv(aji);
            A.val[size_t(q)] = T(0.5) * aji + T(0.5) * conjv(aij);
        }
}

// rows much longer than typical: > max(32, 8 x median length)
template <typename T>
std::vector<uint8_t> long_rows(const Csr<T>& A) {
    const int n = A.nr;
    std::vector<int> len(static_cast<size_t>(n));
    for (int i = 0; i < n; i++) len[i] = int(A.ptr[i + 1] - A.ptr[i]);
    std::vector<int> s(len);
    std::nth_element(s.begin(), s.begin() + n / 2, s.end());
    const int thr = std::max(32, 8 * (n ? s[size_t(n / 2)] : 1));
    std::vector<uint8_t> lng(static_cast<size_t>(n), 0);
    for (int i = 0; i < n; i++) lng[i] = len[i] > thr ? 1 : 0;
    return lng;
}

// With near-null vectors B (n x nb, column-major): a coupling joins an aggregate only if it is
// negative and the two rows of B point the same way (cosine >= 0.9).  Couplings across a seam
// (a rotation maps u to +-v or u to -u) fail this and stay between aggregates, where the
// Galerkin product carries them; an aggregate never straddles a change of chart.
inline bool compatible(double a, const std::vector<double>& B, int nb, int n, int i, int j) {
    if (!(a < 0.0) || !tune().compat) return tune().compat ? false : true;
    double bi = 0.0, bj = 0.0, bij = 0.0;
    for (int c = 0; c < nb; c++) {
        const double x = B[size_t(c) * n + i], y = B[size_t(c) * n + j];
        bi += x * x;
        bj += y * y;
        bij += x * y;
    }
    return bij >= 0.9 * std::sqrt(bi * bj);
}
inline bool compatible(const std::complex<double>&, const std::vector<std::complex<double>>&, int, int, int, int) {
    return true;
}

// Strength graph and greedy aggregation (seed / attach / leftover passes).
template <typename T>
struct Aggregation {
    metriko::HVec<int64_t> sptr;
    metriko::HVec<int> scol;
    metriko::HVec<T> sval;
    std::vector<int> agg, root;
    int nagg = 0;
};

template <typename T>
Aggregation<T> aggregate(const Csr<T>& A, const std::vector<double>& d, double theta,
                         const std::vector<T>* B = nullptr, int nb = 0) {
    Aggregation<T> G;
    const int n = A.nr;
    auto& sptr = G.sptr;
    auto& scol = G.scol;
    auto& sval = G.sval;
    sptr.assign(size_t(n) + 1, 0);
    // long rows (kernel-basis / transition DOFs coupling to a whole seam) stay out of the
    // strength graph: they become singleton aggregates and are not smoothed into others
    std::vector<uint8_t> lng = long_rows(A);
    auto strong = [&](int i, int64_t k) {
        const int j = A.col[k];
        if (j == i || lng[i] || lng[j]) return false;
        if (B && nb > 0 && !compatible(A.val[k], *B, nb, n, i, j)) return false;
        return mag2(A.val[k]) >= theta * theta * std::abs(d[i]) * std::abs(d[j]);
    };
#pragma omp parallel for schedule(static, 4096) if (n > 100000)
    for (int i = 0; i < n; i++) {
        int64_t c = 0;
        for (int64_t k = A.ptr[i]; k < A.ptr[i + 1]; k++) c += strong(i, k) ? 1 : 0;
        sptr[size_t(i) + 1] = c;
    }
    for (int i = 0; i < n; i++) sptr[size_t(i) + 1] += sptr[size_t(i)];
    scol.resize(size_t(sptr[size_t(n)]));
    sval.resize(size_t(sptr[size_t(n)]));
#pragma omp parallel for schedule(static, 4096) if (n > 100000)
    for (int i = 0; i < n; i++) {
        int64_t w = sptr[size_t(i)];
        for (int64_t k = A.ptr[i]; k < A.ptr[i + 1]; k++)
            if (strong(i, k)) { scol[size_t(w)] = A.col[k]; sval[size_t(w)] = A.val[k]; w++; }
    }
    std::vector<int>& agg = G.agg;
    std::vector<int>& root = G.root;
    int& nagg = G.nagg;
    agg.assign(size_t(n), -1);
    for (int i = 0; i < n; i++) {  // seeds: nodes whose strong neighbourhood is untouched
        if (agg[i] != -1) continue;
        bool free_nb = true;
        for (int64_t k = sptr[i]; k < sptr[i + 1]; k++) if (agg[scol[k]] != -1) { free_nb = false; break; }
        if (!free_nb) continue;
        agg[i] = nagg;
        for (int64_t k = sptr[i]; k < sptr[i + 1]; k++) agg[scol[k]] = nagg;
        root.push_back(i);
        nagg++;
    }
    std::vector<int> agg1 = agg;
    const bool attach = tune().attach != 0;
    // attach to the strongest seeded neighbour (each node reads only the seed pass's result)
#pragma omp parallel for schedule(static, 4096) if (n > 100000)
    for (int i = 0; i < n; i++) {
        if (!attach || agg[i] != -1) continue;
        int best = -1;
        double bv = -1.0;
        for (int64_t k = sptr[i]; k < sptr[i + 1]; k++) {
            const int j = scol[k];
            if (agg1[j] >= 0 && mag(sval[k]) > bv) { bv = mag(sval[k]); best = agg1[j]; }
        }
        if (best >= 0) agg[i] = best;
    }
    for (int i = 0; i < n; i++) {  // leftovers
        if (agg[i] != -1) continue;
        agg[i] = nagg;
        for (int64_t k = sptr[i]; k < sptr[i + 1]; k++) if (agg[scol[k]] == -1) agg[scol[k]] = nagg;
        root.push_back(i);
        nagg++;
    }
    return G;
}

// Tentative prolongator with one near-null vector built by BFS inside each aggregate that
// follows the operator's phases (the connection Laplacian's parallel transport).
template <typename T>
Csr<T> tentative(const Csr<T>& A, const std::vector<double>& d, double theta, int& nagg) {
    const int n = A.nr;
    const Aggregation<T> G = aggregate(A, d, theta);
    const auto& sptr = G.sptr;
    const auto& scol = G.scol;
    const auto& sval = G.sval;
    const std::vector<int>& agg = G.agg;
    const std::vector<int>& root = G.root;
    nagg = G.nagg;
    std::vector<T> B(size_t(n), T(0));
    std::vector<uint8_t> seen(size_t(n), 0);
    const bool follow_phase = !(std::is_same<T, double>::value && !tune().phase_real);
    // one BFS per aggregate from its root; a BFS stays inside its aggregate, so the
    // aggregates run in parallel and every node is written by its own aggregate's BFS only
#pragma omp parallel if (n > 100000)
    {
        std::vector<int> q;
#pragma omp for schedule(dynamic, 256)
        for (int a = 0; a < nagg; a++) {
            const int r0 = root[a];
            B[r0] = T(1);
            seen[r0] = 1;
            q.assign(1, r0);
            for (size_t h = 0; h < q.size(); h++) {
                const int u = q[h];
                for (int64_t k = sptr[u]; k < sptr[u + 1]; k++) {
                    const int j = scol[k];
                    if (agg[j] != a || seen[j]) continue;
                    seen[j] = 1;
                    // complex (connection Laplacian): follow the transport phase; real (Laplacian
                    // of the uv coordinates): translations are the near-null space -> constants
                    if (!follow_phase) B[j] = B[u];
                    else B[j] = -conjv(sval[k]) / mag(sval[k]) * B[u];
                    q.push_back(j);
                }
            }
        }
    }
    // members of each aggregate in ascending order: the norms sum in node order
    std::vector<int> mptr(size_t(nagg) + 1, 0), mem(static_cast<size_t>(n));
    for (int i = 0; i < n; i++) mptr[size_t(agg[i]) + 1]++;
    for (int a = 0; a < nagg; a++) mptr[size_t(a) + 1] += mptr[size_t(a)];
    {
        std::vector<int> fill(mptr.begin(), mptr.end() - 1);
        for (int i = 0; i < n; i++) mem[size_t(fill[size_t(agg[i])]++)] = i;
    }
    std::vector<double> nrm(size_t(nagg), 0.0);
#pragma omp parallel for schedule(static, 1024) if (n > 100000)
    for (int a = 0; a < nagg; a++) {
        double s = 0.0;
        for (int q = mptr[size_t(a)]; q < mptr[size_t(a) + 1]; q++) {
            const int i = mem[size_t(q)];
            if (!seen[size_t(i)]) B[size_t(i)] = T(1);
            s += mag2(B[size_t(i)]);
        }
        nrm[size_t(a)] = s;
    }
    Csr<T> P(n, nagg);
    P.col.resize(size_t(n));
    P.val.resize(size_t(n));
#pragma omp parallel for schedule(static, 4096) if (n > 100000)
    for (int i = 0; i < n; i++) {
        P.col[i] = agg[i];
        P.val[i] = B[i] / std::sqrt(nrm[agg[i]]);
        P.ptr[i + 1] = i + 1;
    }
    return P;
}

// Tentative prolongator for a given block of near-null vectors B (n x nb, column-major):
// per aggregate, a thin QR of B's rows (modified Gram-Schmidt, twice, rank-revealing) gives
// the columns of P; R stacked over the aggregates is the coarse level's B.  An aggregate on
// which every vector vanishes keeps one normalised indicator column so no row is orphaned.
template <typename T>
Csr<T> tentative_qr(const Csr<T>& A, const std::vector<double>& d, double theta, const std::vector<T>& B, int nb,
                    std::vector<T>& Bc, int& nc) {
    const int n = A.nr;
    const Aggregation<T> G = aggregate(A, d, theta, &B, nb);
    std::vector<int> mptr(size_t(G.nagg) + 1, 0), mem(static_cast<size_t>(n));
    for (int i = 0; i < n; i++) mptr[G.agg[i] + 1]++;
    for (int a = 0; a < G.nagg; a++) mptr[a + 1] += mptr[a];
    {
        std::vector<int> fill(mptr.begin(), mptr.end() - 1);
        for (int i = 0; i < n; i++) mem[fill[G.agg[i]]++] = i;
    }
    std::vector<double> scale(static_cast<size_t>(nb), 0.0);
    for (int c = 0; c < nb; c++)
        for (int i = 0; i < n; i++) scale[c] = std::max(scale[c], mag(B[size_t(c) * n + i]));
    // per aggregate (independent, in parallel): the orthonormal columns Q (r of them, m rows
    // each, at qoff = mptr[a] * nb) and the coarse rows R = Q^H B (r x nb at roff = a * nb * nb)
    const int nagg = G.nagg;
    metriko::HVec<T> Qall(size_t(n) * size_t(nb)), Rall(size_t(nagg) * size_t(nb) * size_t(nb));
    std::vector<int> rcount(static_cast<size_t>(nagg));
#pragma omp parallel if (n > 100000)
    {
        std::vector<T> v;
#pragma omp for schedule(dynamic, 256)
        for (int a = 0; a < nagg; a++) {
            const int m = mptr[a + 1] - mptr[a];
            const int* rows = mem.data() + mptr[a];
            T* Q = Qall.data() + size_t(mptr[a]) * nb;
            int r = 0;
            for (int c = 0; c < nb; c++) {
                v.assign(static_cast<size_t>(m), T(0));
                for (int i = 0; i < m; i++) v[i] = B[size_t(c) * n + rows[i]];
                double n0 = 0.0;
                for (int i = 0; i < m; i++) n0 += mag2(v[i]);
                n0 = std::sqrt(n0);
                for (int pass = 0; pass < 2; pass++)
                    for (int q = 0; q < r; q++) {
                        T s = T(0);
                        for (int i = 0; i < m; i++) s += metriko::mulv(conjv(Q[size_t(q) * m + i]), v[i]);
                        for (int i = 0; i < m; i++) v[i] -= metriko::mulv(s, Q[size_t(q) * m + i]);
                    }
                double nv = 0.0;
                for (int i = 0; i < m; i++) nv += mag2(v[i]);
                nv = std::sqrt(nv);
                if (!(nv > 1e-10 * std::max(n0, scale[c] * 1e-6)) || nv == 0.0) continue;
                for (int i = 0; i < m; i++) Q[size_t(r) * m + i] = v[i] / nv;
                r++;
            }
            if (r == 0) {
                for (int i = 0; i < m; i++) Q[i] = T(1.0 / std::sqrt(double(m)));
                r = 1;
            }
            T* R = Rall.data() + size_t(a) * nb * nb;
            for (int q = 0; q < r; q++)
                for (int c = 0; c < nb; c++) {
                    T s = T(0);
                    for (int i = 0; i < m; i++) s += metriko::mulv(conjv(Q[size_t(q) * m + i]), B[size_t(c) * n + rows[i]]);
                    R[size_t(q) * nb + c] = s;
                }
            rcount[size_t(a)] = r;
        }
    }
    // coarse numbering: aggregate a owns coarse columns [coff[a], coff[a] + r_a)
    std::vector<int> coff(size_t(nagg) + 1, 0);
    for (int a = 0; a < nagg; a++) coff[size_t(a) + 1] = coff[size_t(a)] + rcount[size_t(a)];
    nc = coff[size_t(nagg)];
    // P: fine row rows[i] of aggregate a holds Q[q][i] in column coff[a] + q (q ascending)
    Csr<T> P(n, nc);
    for (int i = 0; i < n; i++) P.ptr[size_t(i) + 1] = rcount[size_t(G.agg[i])];
    for (int i = 0; i < n; i++) P.ptr[size_t(i) + 1] += P.ptr[size_t(i)];
    P.col.resize(size_t(P.ptr[size_t(n)]));
    P.val.resize(size_t(P.ptr[size_t(n)]));
#pragma omp parallel for schedule(dynamic, 256) if (n > 100000)
    for (int a = 0; a < nagg; a++) {
        const int m = mptr[a + 1] - mptr[a];
        const int* rows = mem.data() + mptr[a];
        const T* Q = Qall.data() + size_t(mptr[a]) * nb;
        for (int i = 0; i < m; i++) {
            int64_t w = P.ptr[size_t(rows[i])];
            for (int q = 0; q < rcount[size_t(a)]; q++, w++) {
                P.col[size_t(w)] = coff[size_t(a)] + q;
                P.val[size_t(w)] = Q[size_t(q) * m + i];
            }
        }
    }
    Bc.assign(size_t(nc) * nb, T(0));
    for (int a = 0; a < nagg; a++)
        for (int q = 0; q < rcount[size_t(a)]; q++)
            for (int c = 0; c < nb; c++)
                Bc[size_t(c) * nc + coff[size_t(a)] + q] = Rall[(size_t(a) * nb + q) * nb + c];
    return P;
}

template <typename T>
void dense_hpd_inverse(const Csr<T>& A, std::vector<double>& out, bool cplx) {
    const int n = A.nr;
    std::vector<T> M(size_t(n) * n, T(0)), L(size_t(n) * n, T(0)), Inv(size_t(n) * n, T(0));
    for (int i = 0; i < n; i++)
        for (int64_t k = A.ptr[i]; k < A.ptr[i + 1]; k++) M[size_t(i) * n + A.col[k]] = A.val[k];
    // Cholesky M = L L^H
    for (int j = 0; j < n; j++) {
        double s = re(M[size_t(j) * n + j]);
        for (int k = 0; k < j; k++) s -= mag2(L[size_t(j) * n + k]);
        if (!(s > 0.0)) s = std::max(1e-300, std::abs(re(M[size_t(j) * n + j])) * 1e-14 + 1e-300);
        const double ljj = std::sqrt(s);
        L[size_t(j) * n + j] = T(ljj);
        for (int i = j + 1; i < n; i++) {
            T t = M[size_t(i) * n + j];
            for (int k = 0; k < j; k++) t -= L[size_t(i) * n + k] * conjv(L[size_t(j) * n + k]);
            L[size_t(i) * n + j] = t / ljj;
        }
    }
    // inverse column by column: L y = e_c, L^H x = y
    std::vector<T> y(static_cast<size_t>(n)), x(static_cast<size_t>(n));
    for (int c = 0; c < n; c++) {
        for (int i = 0; i < n; i++) {
            T t = (i == c) ? T(1) : T(0);
            for (int k = 0; k < i; k++) t -= L[size_t(i) * n + k] * y[k];
            y[i] = t / L[size_t(i) * n + i];
        }
        for (int i = n - 1; i >= 0; i--) {
            T t = y[i];
            for (int k = i + 1; k < n; k++) t -= conjv(L[size_t(k) * n + i]) * x[k];
            x[i] = t / conjv(L[size_t(i) * n + i]);
        }
        for (int i = 0; i < n; i++) Inv[size_t(i) * n + c] = x[i];
    }
    const int vw = cplx ? 2 : 1;
    out.assign(size_t(n) * n * vw, 0.0);
    for (size_t i = 0; i < size_t(n) * n; i++) {
        if constexpr (std::is_same<T, double>::value) out[i] = Inv[i];
        else { out[2 * i] = Inv[i].real(); out[2 * i + 1] = Inv[i].imag(); }
    }
}

// rsrc / cmap: an optional renumbering of the rows / columns (build_sell)
template <typename T>
mk_op* upload(mk_ctx* c, const Csr<T>& A, const mk_solve_opts& o, const int* rsrc = nullptr,
              const int* cmap = nullptr) {
    auto* op = new mk_op();
    try {
        op->ctx = c;
        op->n = A.nr;
        op->m = A.nc;
        op->cplx = !std::is_same<T, double>::value;
        op->nnz = A.nonZeros();
        op->block = o.block_size;
        build_sell(op, A.ptr.data(), A.col.data(), reinterpret_cast<const double*>(A.val.data()), o.slice_height,
                   o.format == MK_FMT_CSR ? 1 : o.sigma, 1.0, rsrc, cmap);
    } catch (...) {
        delete op;
        throw;
    }
    return op;
}

// Host multigrid hierarchy (FP64): level operators, prolongators, restrictions, the
// Chebyshev upper bound per level and the dense inverse of the coarsest operator.
template <typename T>
struct HostHier {
    std::vector<Csr<T>> A, P, R, RR;
    std::vector<double> lmax;
    std::vector<std::vector<double>> dinv;
    std::vector<double> inv;  // coarsest dense inverse (row-major, complex interleaved)
    int nc = 0;
    // per level but the coarsest: the coarse unknown that stands for each node's aggregate
    // (the first column of its tentative prolongator row; -1 for an empty row)
    std::vector<std::vector<int>> key;
};

// Locality order of every level from the aggregates: the coarsest level keeps its order, and
// every finer level lists its nodes by the rank of their aggregate's coarse unknown (nodes of
// one aggregate in index order).  Aggregates are connected neighbourhoods, nested from level to
// level, so rows that share columns end up close together at every scale, and consecutive rows
// of R and P (one level's aggregates in rank order) touch consecutive ranges of the finer
// level.  Returns pos[l][old] = new for every level but the coarsest.
template <typename T>
std::vector<std::vector<int>> hier_order(const HostHier<T>& H) {
    const size_t L = H.key.size();
    std::vector<std::vector<int>> pos(L);
    std::vector<int> rank(size_t(H.A.back().nr));
    std::iota(rank.begin(), rank.end(), 0);
    for (size_t l = L; l-- > 0;) {
        const std::vector<int>& key = H.key[l];
        const int n = int(key.size()), nc = int(rank.size());
        // counting sort by the parent's rank (nodes without a parent last), stable in index
        std::vector<int> cnt(size_t(nc) + 2, 0);
        auto bucket = [&](int i) { return key[size_t(i)] >= 0 ? rank[size_t(key[size_t(i)])] : nc; };
        for (int i = 0; i < n; i++) cnt[size_t(bucket(i)) + 1]++;
        for (int b = 0; b <= nc; b++) cnt[size_t(b) + 1] += cnt[size_t(b)];
        std::vector<int> r(static_cast<size_t>(n));
        for (int i = 0; i < n; i++) r[size_t(i)] = cnt[size_t(bucket(i))]++;
        pos[l] = r;
        rank.swap(r);
    }
    return pos;
}

template <typename T>
HostHier<T> build_host(Csr<T> A, double theta, int coarse_max, const std::vector<T>* B0 = nullptr,
                       int nb = 0) {
    HostHier<T> H;
    const bool cplx = !std::is_same<T, double>::value;
    std::vector<T> B;
    if (B0 && nb > 0) {
        if (B0->size() != size_t(A.nr) * size_t(nb)) throw Error(MK_ERR_ARG, "multigrid: near-null block size mismatch");
        B = *B0;
    } else {
        nb = 0;
    }
    const int maxLevels = 14;
    const bool verbose = std::getenv("METRIKO_VERBOSE") != nullptr;
    double t_lap = 0.0;
    auto lap = [&](const char* what) {
        if (!verbose) return;
        const double t = std::chrono::duration<double, std::milli>(
                             std::chrono::steady_clock::now().time_since_epoch()).count();
        if (t_lap > 0.0) std::fprintf(stderr, "[mg]   %s %.0f ms\n", what, t - t_lap);
        t_lap = t;
    };
    for (int l = 0; l < maxLevels; l++) {
        lap("level start");
        std::vector<double> d = diag_of(A);
        if (l == 0)
            for (int i = 0; i < A.nr; i++)
                if (!(d[i] > 0.0)) throw Error(MK_ERR_NOT_SPD, "multigrid: non-positive diagonal (matrix not SPD/HPD)");
        // smoothing diagonal: d_i, raised to half the absolute row sum where the row is not
        // diagonally dominant, so every Gershgorin disc of Ds^-1 A lies in [0, 2] and a few
        // strongly coupled rows cannot stretch the Chebyshev interval of all the others
        std::vector<double> ds(d);
        if (tune().l1) {
#pragma omp parallel for schedule(static, 4096) if (A.nr > 100000)
            for (int i = 0; i < A.nr; i++) {
                double sa = 0.0;
                for (int64_t k = A.ptr[i]; k < A.ptr[i + 1]; k++) sa += mag(A.val[k]);
                ds[i] = std::max(d[i], 0.5 * sa);
            }
        }
        lap("diagonals");
        const double lmax = tune().lmax_fac * rho_dinv_a(A, ds);
        lap("spectral bound");
        std::vector<double> di(d.size());
        for (size_t i = 0; i < d.size(); i++) di[i] = ds[i] != 0.0 ? 1.0 / ds[i] : 1.0;
        if (verbose) {
            static double t_prev = 0.0;
            const double t_now = std::chrono::duration<double, std::milli>(
                                     std::chrono::steady_clock::now().time_since_epoch()).count();
            if (l == 0) t_prev = t_now;
            std::fprintf(stderr, "[mg] +%.0f ms\n", t_now - t_prev);
            t_prev = t_now;
            int64_t mx = 0;
            for (int i = 0; i < A.nr; i++) mx = std::max<int64_t>(mx, A.ptr[i + 1] - A.ptr[i]);
            std::fprintf(stderr, "[mg] level %d: n=%d nnz=%lld maxrow=%lld lmax=%.3g\n", l, A.nr,
                         (long long)A.nonZeros(), (long long)mx, lmax);
        }
        int nagg = 0;
        Csr<T> Pt;
        bool last = A.nr <= coarse_max || l == maxLevels - 1;
        if (!last) {
            std::vector<T> Bc;
            if (nb > 0) Pt = tentative_qr(A, d, theta, B, nb, Bc, nagg);
            else Pt = tentative(A, d, theta, nagg);
            {
                std::vector<int> key(static_cast<size_t>(A.nr));
#pragma omp parallel for schedule(static, 65536) if (A.nr > 200000)
                for (int i = 0; i < A.nr; i++) key[size_t(i)] = Pt.ptr[i + 1] > Pt.ptr[i] ? Pt.col[Pt.ptr[i]] : -1;
                H.key.push_back(std::move(key));
            }
            lap("aggregation + tentative");
            if (nb > 0) B = std::move(Bc);
            if (nagg >= A.nr * 0.9 || nagg < 1) {  // no coarsening possible: make this the coarsest level
                if (A.nr > 4000) throw Error(MK_ERR_UNSUPPORTED, "multigrid: aggregation stalled on a large level");
                last = true;
                H.key.pop_back();
            }
        }
        H.lmax.push_back(lmax);
        H.dinv.push_back(std::move(di));
        if (last) {
            dense_hpd_inverse(A, H.inv, cplx);
            H.nc = A.nr;
            H.A.push_back(std::move(A));
            break;
        }
        const double rho = lmax / tune().lmax_fac;
            // With exact near-null vectors the ordinary rows see their couplings to long rows
            // (so P B_c = B holds there too); long rows themselves are never smoothed: a smoothed
            // long row would couple every aggregate along its seam in R A P.
            const bool keep_long = nb > 0 && tune().smooth_long;
            const double omega = 4.0 / (3.0 * rho);
            // P = Pt - omega D^-1 Af Pt, Af = A without the couplings of long rows (and, without
            // exact near-null vectors, of long columns)
            std::vector<uint8_t> lng = long_rows(A);
            auto keep = [&](int i, int j) { return j == i || !(lng[i] || (lng[j] && !keep_long)); };
            // without long rows Af is A itself
            const bool any_long = std::find(lng.begin(), lng.end(), uint8_t(1)) != lng.end();
            Csr<T> Af(any_long ? A.nr : 0, any_long ? A.nc : 0);
            if (any_long) {
                auto& np = Af.ptr;
#pragma omp parallel for schedule(static, 4096)
                for (int i = 0; i < A.nr; i++) {
                    int64_t c = 0;
                    for (int64_t k = A.ptr[i]; k < A.ptr[i + 1]; k++) c += keep(i, A.col[k]) ? 1 : 0;
                    np[size_t(i) + 1] = c;
                }
                for (int i = 0; i < A.nr; i++) np[size_t(i) + 1] += np[size_t(i)];
                Af.col.resize(size_t(np[size_t(A.nr)]));
                Af.val.resize(size_t(np[size_t(A.nr)]));
#pragma omp parallel for schedule(static, 4096)
                for (int i = 0; i < A.nr; i++) {
                    int64_t w = np[size_t(i)];
                    for (int64_t k = A.ptr[i]; k < A.ptr[i + 1]; k++)
                        if (keep(i, A.col[k])) { Af.col[size_t(w)] = A.col[k]; Af.val[size_t(w)] = A.val[k]; w++; }
                }
            }
            lap("filtered operator");
            Csr<T> AP = metriko::spgemm(any_long ? Af : A, Pt);
            lap("A Pt");
            // P = Pt + rows of A Pt scaled by -omega / ds_i (long rows: Pt's row alone), each row
            // the merge of the two sorted rows: an entry in both is Pt + s, s = (-omega / ds_i) (A Pt)
            Csr<T> P(A.nr, Pt.nc);
            {
                auto merge = [&](int i, int64_t w, bool fill) {
                    int64_t p = Pt.ptr[i], q = lng[i] ? AP.ptr[i + 1] : AP.ptr[i], n = 0;
                    const int64_t pe = Pt.ptr[i + 1], qe = AP.ptr[i + 1];
                    const double di = ds[i] != 0.0 ? ds[i] : 1.0;
                    while (p < pe || q < qe) {
                        if (q >= qe || (p < pe && Pt.col[p] < AP.col[q])) {
                            if (fill) { P.col[size_t(w)] = Pt.col[p]; P.val[size_t(w)] = Pt.val[p]; }
                            p++;
                        } else if (p >= pe || AP.col[q] < Pt.col[p]) {
                            if (fill) {
                                const T sv = -omega / di * AP.val[q];
                                P.col[size_t(w)] = AP.col[q];
                                P.val[size_t(w)] = sv;
                            }
                            q++;
                        } else {
                            if (fill) {
                                const T sv = -omega / di * AP.val[q];
                                P.col[size_t(w)] = Pt.col[p];
                                P.val[size_t(w)] = Pt.val[p] + sv;
                            }
                            p++;
                            q++;
                        }
                        w++;
                        n++;
                    }
                    return n;
                };
#pragma omp parallel for schedule(static, 4096)
                for (int i = 0; i < A.nr; i++) P.ptr[size_t(i) + 1] = merge(i, 0, false);
                for (int i = 0; i < A.nr; i++) P.ptr[size_t(i) + 1] += P.ptr[size_t(i)];
                P.col.resize(size_t(P.ptr[size_t(A.nr)]));
                P.val.resize(size_t(P.ptr[size_t(A.nr)]));
#pragma omp parallel for schedule(static, 4096)
                for (int i = 0; i < A.nr; i++) merge(i, P.ptr[size_t(i)], true);
            }
            lap("smoothed P");
            Csr<T> R = conj_transpose(P);
            lap("R");
            Csr<T> AP2 = metriko::spgemm(A, P);
            Csr<T> Ac = metriko::spgemm(R, AP2);
            lap("R A P");
            // [R | -R A] with R A = (A P)^H (A Hermitian): the residual restriction operator,
            // built on request only (its transpose costs more host time than the fused
            // restriction saves over the 50 inverse-iteration solves)
            if (tune().rr) {
                Csr<T> RA = conj_transpose(AP2);
                Csr<T> RR(R.nr, 2 * A.nc);
                for (int i = 0; i < R.nr; i++)
                    RR.ptr[size_t(i) + 1] = RR.ptr[size_t(i)] + (R.ptr[i + 1] - R.ptr[i]) + (RA.ptr[i + 1] - RA.ptr[i]);
                RR.col.resize(size_t(RR.ptr[size_t(R.nr)]));
                RR.val.resize(size_t(RR.ptr[size_t(R.nr)]));
#pragma omp parallel for schedule(static, 1024)
                for (int i = 0; i < R.nr; i++) {
                    int64_t w = RR.ptr[size_t(i)];
                    for (int64_t k = R.ptr[i]; k < R.ptr[i + 1]; k++, w++) { RR.col[size_t(w)] = R.col[k]; RR.val[size_t(w)] = R.val[k]; }
                    for (int64_t k = RA.ptr[i]; k < RA.ptr[i + 1]; k++, w++) {
                        RR.col[size_t(w)] = A.nc + RA.col[k];
                        RR.val[size_t(w)] = -RA.val[k];
                    }
                }
                H.RR.push_back(std::move(RR));
                lap("residual restriction");
            }
            // exact Hermitian symmetry of the coarse operator
            hermitian_average(Ac);
            H.A.push_back(std::move(A));
            H.P.push_back(std::move(P));
            H.R.push_back(std::move(R));
            A = std::move(Ac);
    }
    return H;
}

// order_out: when given (and the hierarchy has more than one level), every level but the
// coarsest is stored in the locality order of hier_order -- its operator's rows and columns,
// P's rows and R's columns (and the next level's order for P's columns and R's rows) and the
// smoother's diagonal -- and the finest level's pos[old] = new is returned there; the caller's
// vectors for this hierarchy are then in that order.  Each renumbered row keeps its entries'
// order, so every product computes the same numbers as without the renumbering.
template <typename T>
Mg* build_impl(mk_ctx* c, Csr<T>&& A0, const mk_solve_opts& o, double theta, int coarse_max,
               const std::vector<T>* nullvec = nullptr, int nnull = 0, std::vector<int>* order_out = nullptr) {
    HostHier<T> H = build_host<T>(std::move(A0), theta, coarse_max, nullvec, nnull);
    const auto t_up = std::chrono::steady_clock::now();
    // per level but the coarsest: pos (stored index of each unknown) and its inverse
    std::vector<std::vector<int>> pos, inv;
    if (order_out && H.A.size() > 1) {
        pos = hier_order(H);
        inv.resize(pos.size());
        for (size_t l = 0; l < pos.size(); l++) {
            inv[l].resize(pos[l].size());
            for (size_t i = 0; i < pos[l].size(); i++) inv[l][size_t(pos[l][i])] = int(i);
        }
    }
    const bool verbose = std::getenv("METRIKO_VERBOSE") != nullptr;
    auto since_up = [&]() {
        return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t_up).count();
    };
    if (verbose) std::fprintf(stderr, "[mg]   order %.0f ms\n", since_up());
    auto P_ = [&](size_t l) -> const int* { return l < pos.size() ? pos[l].data() : nullptr; };
    auto I_ = [&](size_t l) -> const int* { return l < inv.size() ? inv[l].data() : nullptr; };
    auto* g = new Mg();
    try {
        g->cplx = !std::is_same<T, double>::value;
        g->deg = tune().deg;
        for (size_t l = 0; l < H.A.size(); l++) {
            MgLevel L;
            L.n = H.A[l].nr;
            L.lmax = H.lmax[l];
            L.A = upload(c, H.A[l], o, I_(l), P_(l));
            {  // the smoother's inverse diagonal replaces the plain one
                std::vector<double> dv = H.dinv[l];
                if (P_(l))
                    for (size_t i = 0; i < dv.size(); i++) dv[size_t(pos[l][i])] = H.dinv[l][i];
                std::vector<float> dv32(dv.begin(), dv.end());
                L.A->dinv64.upload(dv.data(), dv.size(), c->stream);
                L.A->dinv32.upload(dv32.data(), dv32.size(), c->stream);
                MK_HIP(hipStreamSynchronize(c->stream));
            }
            g->nnz_total += H.A[l].nonZeros();
            if (l < H.P.size()) {
                L.P = upload(c, H.P[l], o, I_(l), P_(l + 1));
                L.R = upload(c, H.R[l], o, I_(l + 1), P_(l));
                if (l < H.RR.size()) {
                    // [R | -R A]: columns [0, n) and [n, 2n) are both this level's unknowns
                    std::vector<int> pos2;
                    if (P_(l)) {
                        const size_t n = pos[l].size();
                        pos2.resize(2 * n);
                        for (size_t i = 0; i < n; i++) { pos2[i] = pos[l][i]; pos2[i + n] = pos[l][i] + int(n); }
                    }
                    L.RR = upload(c, H.RR[l], o, I_(l + 1), pos2.empty() ? nullptr : pos2.data());
                }
            }
            g->lv.push_back(std::move(L));
            if (verbose) std::fprintf(stderr, "[mg]   level %zu uploaded %.0f ms\n", l, since_up());
        }
        g->nc = H.nc;
        g->inv64.upload(H.inv.data(), H.inv.size(), c->stream);
        std::vector<float> inv32(H.inv.begin(), H.inv.end());
        g->inv32.upload(inv32.data(), inv32.size(), c->stream);
        MK_HIP(hipStreamSynchronize(c->stream));
        if (verbose) std::fprintf(stderr, "[mg]   coarsest inverse uploaded %.0f ms\n", since_up());
        for (auto& L : g->lv) {
            const size_t len = size_t(L.n) * (g->cplx ? 2 : 1);
            L.b64.alloc(len); L.x64.alloc(len); L.r64.alloc(len); L.d64.alloc(len); L.y64.alloc(len);
            L.b32.alloc(len); L.x32.alloc(len); L.r32.alloc(len); L.d32.alloc(len); L.y32.alloc(len);
        }
        MK_LOG("[mg] hierarchy upload %.0f ms\n",
               std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t_up).count());
    } catch (...) {
        mg_free(g);
        throw;
    }
    if (order_out) {
        order_out->clear();
        if (!pos.empty()) order_out->swap(pos[0]);
    }
    return g;
}

// ---- device kernels ------------------------------------------------------------------------
// x = c0 * dinv .* r (zero start) or x += ...; d = same increment.  Vectors of n elements (x2 complex)
template <typename T, bool CPLX>
__global__ void k_cheb_first(int n, const T* __restrict__ dinv, const T* __restrict__ r, T* __restrict__ d,
                             T* __restrict__ x, T c0, int zero_start, const int* skip) {
    if (skip && *skip) return;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const T s = c0 * dinv[i];
    if constexpr (CPLX) {
        const T d0 = s * r[2 * i], d1 = s * r[2 * i + 1];
        d[2 * i] = d0; d[2 * i + 1] = d1;
        if (zero_start) { x[2 * i] = d0; x[2 * i + 1] = d1; }
        else { x[2 * i] += d0; x[2 * i + 1] += d1; }
    } else {
        const T d0 = s * r[i];
        d[i] = d0;
        if (zero_start) x[i] = d0;
        else x[i] += d0;
    }
}
// x = Inv b (dense n x n, row-major): one wavefront per row, lanes stride the row (coalesced
// reads of Inv), shuffle reduction; 4 rows per 256-thread block
template <typename T, bool CPLX>
__global__ void __launch_bounds__(256) k_dense_gemv(int n, const T* __restrict__ inv, const T* __restrict__ b,
                                                    T* __restrict__ x, const int* skip) {
    if (skip && *skip) return;
    const int i = blockIdx.x * 4 + int(threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (i >= n) return;  // uniform inside the wavefront
    T ar = 0, ai = 0;
    for (int j = lane; j < n; j += 64) {
        if constexpr (CPLX) {
            const T mr = inv[2 * (size_t(i) * n + j)], mi = inv[2 * (size_t(i) * n + j) + 1];
            const T br = b[2 * j], bi = b[2 * j + 1];
            ar += mr * br - mi * bi;
            ai += mr * bi + mi * br;
        } else {
            ar += inv[size_t(i) * n + j] * b[j];
        }
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        ar += __shfl_xor(ar, off, 64);
        if constexpr (CPLX) ai += __shfl_xor(ai, off, 64);
    }
    if (lane == 0) {
        if constexpr (CPLX) { x[2 * i] = ar; x[2 * i + 1] = ai; }
        else x[i] = ar;
    }
}

inline unsigned nb(size_t n) { return unsigned((n + 255) / 256); }

template <typename T>
struct Sel;
template <>
struct Sel<double> {
    static const double* val(const mk_op* A) { return A->val64.p; }
    static const double* dinv(const mk_op* A) { return A->dinv64.p; }
    static double* b(MgLevel& L) { return L.b64.p; }
    static double* x(MgLevel& L) { return L.x64.p; }
    static double* r(MgLevel& L) { return L.r64.p; }
    static double* d(MgLevel& L) { return L.d64.p; }
    static double* y(MgLevel& L) { return L.y64.p; }
    static const double* inv(const Mg* g) { return g->inv64.p; }
    static double bytes(const mk_op* A) { return A->bytes64; }
};
template <>
struct Sel<float> {
    static const float* val(const mk_op* A) { return A->val32.p; }
    static const float* dinv(const mk_op* A) { return A->dinv32.p; }
    static float* b(MgLevel& L) { return L.b32.p; }
    static float* x(MgLevel& L) { return L.x32.p; }
    static float* r(MgLevel& L) { return L.r32.p; }
    static float* d(MgLevel& L) { return L.d32.p; }
    static float* y(MgLevel& L) { return L.y32.p; }
    static const float* inv(const Mg* g) { return g->inv32.p; }
    static double bytes(const mk_op* A) { return A->bytes32; }
};

// Chebyshev smoothing of A x = b (degree g->deg, interval [lmax/lmin_div, lmax] of D^-1 A).
// Each step after the first of a zero start is one fused kernel (residual, direction and
// update of a row together) that reads one iterate and writes the other (x <-> L.y), so
// every step is a single launch; the buffers are chosen so the result lands in x.
template <typename T>
void chebyshev(mk_ctx* c, const Mg* g, MgLevel& L, const T* b, T* x, bool zero_start, const int* skip) {
    const double lmax = L.lmax, lmin = L.lmax / tune().lmin_div;
    const double theta = 0.5 * (lmax + lmin), delta = 0.5 * (lmax - lmin), sigma = theta / delta;
    double rho0 = 1.0 / sigma;
    T* y = Sel<T>::y(L);
    T* d = Sel<T>::d(L);
    const int n = L.n;
    // fused residual steps after the start; a zero start of degree >= 2 does its first two
    // steps in one launch (cheb0) and continues from there
    int steps = zero_start ? g->deg - 1 : g->deg;
    T* cur = x;
    if (zero_start && g->deg >= 2) {
        steps = g->deg - 2;
        cur = (steps % 2 == 0) ? x : y;
        const double rho1 = 1.0 / (2.0 * sigma - rho0);
        cheb0_launch<T>(c, L.A, b, cur, d, T(1.0 / theta), T(rho1 * rho0), T(2.0 * rho1 / delta), skip);
        rho0 = rho1;
    } else if (zero_start) {
        cur = (steps % 2 == 0) ? x : y;
        if (g->cplx)
            hipLaunchKernelGGL((k_cheb_first<T, true>), dim3(nb(n)), dim3(256), 0, c->stream, n, Sel<T>::dinv(L.A), b,
                               d, cur, T(1.0 / theta), 1, skip);
        else
            hipLaunchKernelGGL((k_cheb_first<T, false>), dim3(nb(n)), dim3(256), 0, c->stream, n, Sel<T>::dinv(L.A),
                               b, d, cur, T(1.0 / theta), 1, skip);
    }
    for (int k = 0; k < steps; k++) {
        T* nxt = (cur == x) ? y : x;
        double a = 0.0, sc;
        int first = 0;
        if (!zero_start && k == 0) {  // first step of a warm start: d = D^-1 r / theta
            first = 1;
            sc = 1.0 / theta;
        } else {
            const double rho1 = 1.0 / (2.0 * sigma - rho0);
            a = rho1 * rho0;
            sc = 2.0 * rho1 / delta;
            rho0 = rho1;
        }
        cheb_launch<T>(c, L.A, cur, nxt, b, d, T(a), T(sc), first, skip);
        cur = nxt;
    }
    if (cur != x) copy_launch<T>(c, size_t(n) * (g->cplx ? 2 : 1), cur, x, skip);
}

template <typename T>
void vcycle(mk_ctx* c, Mg* g, int l, const T* b, T* x, const int* skip) {
    MgLevel& L = g->lv[size_t(l)];
    if (l + 1 == int(g->lv.size())) {
        if (g->cplx)
            hipLaunchKernelGGL((k_dense_gemv<T, true>), dim3(unsigned((g->nc + 3) / 4)), dim3(256), 0, c->stream, g->nc,
                               Sel<T>::inv(g), b, x, skip);
        else
            hipLaunchKernelGGL((k_dense_gemv<T, false>), dim3(unsigned((g->nc + 3) / 4)), dim3(256), 0, c->stream, g->nc,
                               Sel<T>::inv(g), b, x, skip);
        return;
    }
    chebyshev<T>(c, g, L, b, x, true, skip);
    MgLevel& Cn = g->lv[size_t(l) + 1];
    if (L.RR) {  // b_c = [R | -R A] [b; x] = R (b - A x) in one launch
        spmv2_launch<T>(c, L.RR, L.n, b, x, Sel<T>::b(Cn), skip);
    } else {
        T* r = Sel<T>::r(L);
        spmv_launch<T>(c, L.A, Sel<T>::val(L.A), x, r, b, SPMV_RESID_NORED, nullptr, Sel<T>::bytes(L.A), skip);
        spmv_launch<T>(c, L.R, Sel<T>::val(L.R), r, Sel<T>::b(Cn), nullptr, SPMV_PLAIN, nullptr, Sel<T>::bytes(L.R),
                       skip);
    }
    vcycle<T>(c, g, l + 1, Sel<T>::b(Cn), Sel<T>::x(Cn), skip);
    spmv_launch<T>(c, L.P, Sel<T>::val(L.P), Sel<T>::x(Cn), x, nullptr, SPMV_ACCUM, nullptr, Sel<T>::bytes(L.P), skip);
    chebyshev<T>(c, g, L, b, x, false, skip);
}

// ---- batched V-cycle (n x mw row-major blocks, FP32 or FP64) --------------------------------------
// x = c0 D^-1 r: the first Chebyshev iterate from a zero start
template <typename T>
__global__ void k_bcheb_first(size_t len, int sh, const T* __restrict__ dinv, const T* __restrict__ r,
                              T* __restrict__ x, T c0, const int* skip) {
    if (skip && *skip) return;
    const size_t e = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (e >= len) return;
    x[e] = c0 * dinv[e >> sh] * r[e];
}
// X = Inv B on the coarsest level, one thread per (row, column)
template <typename T>
__global__ void k_bdense(int nc, int sh, const T* __restrict__ inv, const T* __restrict__ b, T* __restrict__ x,
                         const int* skip) {
    if (skip && *skip) return;
    const size_t e = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    const size_t mw = size_t(1) << sh;
    if (e >= size_t(nc) * mw) return;
    const size_t i = e >> sh, j = e & (mw - 1);
    T a = T(0);
    for (int k = 0; k < nc; k++) a += inv[i * size_t(nc) + k] * b[size_t(k) * mw + j];
    x[e] = a;
}

// the block buffers of a level in precision T
template <typename T>
struct BlockBufs { T *b, *x, *r, *y; };
template <typename T>
static BlockBufs<T> block_bufs(MgLevel& L) {
    if constexpr (std::is_same<T, double>::value) return {L.bb64.p, L.bx64.p, L.br64.p, L.by64.p};
    else return {L.bb32.p, L.bx32.p, L.br32.p, L.by32.p};
}

// Chebyshev(deg) smoothing of the block in the three-term form.  Every step is one fused kernel
// (residual and update) that writes over the buffer holding the iterate before the one it
// reads; the two buffers x and y alternate, the first chosen so that the last step lands in x.
// After a zero start the first iterate x_1 = c0 D^-1 b is never stored: the second step forms
// it inside its gathers (mode 8) and the third from the row's own b (mode 9).  With deg = 1
// x_1 is the result and is written (k_bcheb_first).
template <typename T>
static void chebyshev_block(mk_ctx* c, const Mg* g, MgLevel& L, int mw, int sh, const T* b, T* x, bool zero_start,
                            const int* skip) {
    const double lmax = L.lmax, lmin = L.lmax / tune().lmin_div;
    const double theta = 0.5 * (lmax + lmin), delta = 0.5 * (lmax - lmin), sigma = theta / delta;
    double rho0 = 1.0 / sigma;
    const size_t len = size_t(L.n) * mw;
    T* y = block_bufs<T>(L).y;
    const T* dinv;
    if constexpr (std::is_same<T, double>::value) dinv = L.A->dinv64.p;
    else dinv = L.A->dinv32.p;
    T* cur = x;
    int fused = g->deg;
    const T c0 = T(1.0 / theta);
    if (zero_start && g->deg == 1) {
        hipLaunchKernelGGL((k_bcheb_first<T>), dim3(nb(len)), dim3(256), 0, c->stream, len, sh, dinv, b, x, c0,
                           skip);
        return;
    }
    if (zero_start) {
        // deg - 1 fused steps after the implicit x_1: the first writes the buffer that makes
        // the last one land in x (cur only selects it; mode 8 does not read it)
        cur = ((g->deg - 1) % 2 == 0) ? x : y;
        fused = g->deg - 1;
    }
    for (int k = 0; k < fused; k++) {
        T* out = (cur == x) ? y : x;
        if (!zero_start && k == 0) {
            spmm_cheb_launch<T>(c, L.A, mw, cur, out, b, T(0), c0, 5, skip);
        } else {
            const double rho1 = 1.0 / (2.0 * sigma - rho0);
            const int mode = !zero_start ? 6 : k == 0 ? 8 : k == 1 ? 9 : 6;
            spmm_cheb_launch<T>(c, L.A, mw, cur, out, b, T(rho1 * rho0), T(2.0 * rho1 / delta), mode, skip, c0);
            rho0 = rho1;
        }
        cur = out;
    }
    if (cur != x) copy_launch<T>(c, len, cur, x, skip);
}

template <typename T>
static void vcycle_block(mk_ctx* c, Mg* g, int l, int mw, int sh, const T* b, T* x, const int* skip) {
    MgLevel& L = g->lv[size_t(l)];
    if (l + 1 == int(g->lv.size())) {
        const T* inv;
        if constexpr (std::is_same<T, double>::value) inv = g->inv64.p;
        else inv = g->inv32.p;
        hipLaunchKernelGGL((k_bdense<T>), dim3(nb(size_t(g->nc) * mw)), dim3(256), 0, c->stream, g->nc, sh, inv, b, x,
                           skip);
        return;
    }
    const BlockBufs<T> Lb = block_bufs<T>(L);
    chebyshev_block<T>(c, g, L, mw, sh, b, x, true, skip);
    spmm_launch<T>(c, L.A, mw, x, Lb.r, b, SPMV_RESID_NORED, skip);
    MgLevel& Cn = g->lv[size_t(l) + 1];
    const BlockBufs<T> Cb = block_bufs<T>(Cn);
    spmm_launch<T>(c, L.R, mw, Lb.r, Cb.b, nullptr, SPMV_PLAIN, skip);
    vcycle_block<T>(c, g, l + 1, mw, sh, Cb.b, Cb.x, skip);
    spmm_launch<T>(c, L.P, mw, Cb.x, x, nullptr, SPMV_ACCUM, skip);
    chebyshev_block<T>(c, g, L, mw, sh, b, x, false, skip);
}

// ---- host V-cycle / PCG (test and tuning aid; same algorithm as the device V-cycle) ----------
template <typename T>
void host_cheb(const Csr<T>& A, const std::vector<double>& dinv, double lmax, int deg, const std::vector<T>& b,
               std::vector<T>& x, bool zero) {
    const size_t n = b.size();
    const double lmin = lmax / tune().lmin_div, theta = 0.5 * (lmax + lmin), delta = 0.5 * (lmax - lmin),
                 sigma = theta / delta;
    double rho0 = 1.0 / sigma;
    std::vector<T> r(n), d(n), ax(n);
    if (zero) { r = b; x.assign(n, T(0)); }
    else { A.multiply(x.data(), ax.data()); for (size_t i = 0; i < n; i++) r[i] = b[i] - ax[i]; }
    for (size_t i = 0; i < n; i++) { d[i] = (1.0 / theta) * dinv[i] * r[i]; x[i] += d[i]; }
    for (int k = 1; k < deg; k++) {
        const double rho1 = 1.0 / (2.0 * sigma - rho0);
        A.multiply(x.data(), ax.data());
        for (size_t i = 0; i < n; i++) {
            d[i] = (rho1 * rho0) * d[i] + (2.0 * rho1 / delta) * dinv[i] * (b[i] - ax[i]);
            x[i] += d[i];
        }
        rho0 = rho1;
    }
}

template <typename T>
void host_vcycle(const HostHier<T>& H, size_t l, const std::vector<T>& b, std::vector<T>& x) {
    const size_t n = b.size();
    if (l + 1 == H.A.size()) {
        x.assign(n, T(0));
        for (size_t i = 0; i < n; i++) {
            T s = T(0);
            for (size_t j = 0; j < n; j++) {
                if constexpr (std::is_same<T, double>::value) s += H.inv[i * n + j] * b[j];
                else s += cd(H.inv[2 * (i * n + j)], H.inv[2 * (i * n + j) + 1]) * b[j];
            }
            x[i] = s;
        }
        return;
    }
    host_cheb(H.A[l], H.dinv[l], H.lmax[l], tune().deg, b, x, true);
    std::vector<T> ax(n), r(n);
    H.A[l].multiply(x.data(), ax.data());
    for (size_t i = 0; i < n; i++) r[i] = b[i] - ax[i];
    std::vector<T> bc(size_t(H.R[l].nr)), xc;
    H.R[l].multiply(r.data(), bc.data());
    host_vcycle(H, l + 1, bc, xc);
    std::vector<T> px(n);
    H.P[l].multiply(xc.data(), px.data());
    for (size_t i = 0; i < n; i++) x[i] += px[i];
    host_cheb(H.A[l], H.dinv[l], H.lmax[l], tune().deg, b, x, false);
}

inline double redot(double a, double b) { return a * b; }
inline double redot(const cd& a, const cd& b) { return (std::conj(a) * b).real(); }

template <typename T>
int host_pcg(const Csr<T>& A, const HostHier<T>& H, const T* b, T* x, double tol, int maxit, double* relres) {
    const size_t n = size_t(A.nr);
    if (maxit == 0) {  // one V-cycle: x = M^-1 b
        std::vector<T> bv(b, b + n), xv;
        host_vcycle(H, 0, bv, xv);
        std::copy(xv.begin(), xv.end(), x);
        return 0;
    }
    std::vector<T> r(n), z, p, q(n), bb(b, b + n);
    A.multiply(x, q.data());
    double bn = 0.0;
    for (size_t i = 0; i < n; i++) { r[i] = b[i] - q[i]; bn += redot(b[i], b[i]); }
    host_vcycle(H, 0, r, z);
    p = z;
    double rz = 0.0, rr = 0.0;
    for (size_t i = 0; i < n; i++) { rz += redot(r[i], z[i]); rr += redot(r[i], r[i]); }
    int it = 0;
    while (it < maxit && rr > tol * tol * bn) {
        A.multiply(p.data(), q.data());
        double pq = 0.0;
        for (size_t i = 0; i < n; i++) pq += redot(p[i], q[i]);
        const double alpha = rz / pq;
        rr = 0.0;
        for (size_t i = 0; i < n; i++) { x[i] += alpha * p[i]; r[i] -= alpha * q[i]; rr += redot(r[i], r[i]); }
        it++;
        if (rr <= tol * tol * bn) break;
        host_vcycle(H, 0, r, z);
        double rz2 = 0.0;
        for (size_t i = 0; i < n; i++) rz2 += redot(r[i], z[i]);
        const double beta = rz2 / rz;
        rz = rz2;
        for (size_t i = 0; i < n; i++) p[i] = z[i] + beta * p[i];
    }
    if (relres) *relres = std::sqrt(rr / (bn > 0 ? bn : 1.0));
    return it;
}

}  // namespace

// Host-only multigrid PCG (no GPU): builds the same hierarchy the device uses and runs PCG
// with a host V-cycle; returns iterations, levels and the final relative residual.
extern "C" mk_status mk_mg_host_pcg(int n, const int64_t* ptr, const int32_t* col, const double* val, int cplx,
                                    const double* nullvec, int nnull, const double* b, double* x, double tol,
                                    int maxit, int* iters, int* levels, double* relres) {
    if (n <= 0 || !ptr || !col || !val || !b || !x || nnull < 0 || (nnull > 0 && !nullvec)) return MK_ERR_ARG;
    try {
        if (cplx) {
            Csr<cd> A(n, n);
            A.ptr.assign(ptr, ptr + n + 1);
            A.col.assign(col, col + ptr[n]);
            A.val.resize(size_t(ptr[n]));
            for (int64_t k = 0; k < ptr[n]; k++) A.val[k] = cd(val[2 * k], val[2 * k + 1]);
            std::vector<cd> B;
            for (size_t q = 0; q < size_t(n) * size_t(nnull); q++) B.emplace_back(nullvec[2 * q], nullvec[2 * q + 1]);
            HostHier<cd> H = build_host<cd>(Csr<cd>(A), tune().theta_cplx, tune().coarse_cplx, nnull ? &B : nullptr, nnull);
            if (levels) *levels = int(H.A.size());
            int it = host_pcg<cd>(A, H, reinterpret_cast<const cd*>(b), reinterpret_cast<cd*>(x), tol, maxit, relres);
            if (iters) *iters = it;
        } else {
            Csr<double> A(n, n);
            A.ptr.assign(ptr, ptr + n + 1);
            A.col.assign(col, col + ptr[n]);
            A.val.assign(val, val + ptr[n]);
            std::vector<double> B;
            if (nnull > 0) B.assign(nullvec, nullvec + size_t(n) * nnull);
            HostHier<double> H = build_host<double>(Csr<double>(A), tune().theta_real, tune().coarse_real,
                                                    nnull ? &B : nullptr, nnull);
            if (levels) *levels = int(H.A.size());
            int it = host_pcg<double>(A, H, b, x, tol, maxit, relres);
            if (iters) *iters = it;
        }
        return MK_OK;
    } catch (const Error& e) {
        return e.code;
    } catch (const std::exception&) {
        return MK_ERR_ARG;
    }
}

namespace {
}  // namespace

Mg* mg_build_real(mk_ctx* c, int n, const int64_t* ptr, const int* col, const double* val, const mk_solve_opts& o,
                  const std::vector<double>* nullvec, int nnull, std::vector<int>* order_out) {
    Csr<double> A(n, n);
    A.ptr.assign(ptr, ptr + n + 1);
    A.col.resize(size_t(ptr[n]));
    A.val.resize(size_t(ptr[n]));
#pragma omp parallel for schedule(static) if (ptr[n] > 200000)
    for (int64_t k = 0; k < ptr[n]; k++) {
        A.col[size_t(k)] = col[k];
        A.val[size_t(k)] = val[k];
    }
    Mg* g = build_impl<double>(c, std::move(A), o, tune().theta_real, tune().coarse_real, nullvec, nnull, order_out);
    return g;
}
Mg* mg_build_complex(mk_ctx* c, int n, const int64_t* ptr, const int* col, const double* val,
                     const mk_solve_opts& o) {
    Csr<cd> A(n, n);
    A.ptr.assign(ptr, ptr + n + 1);
    A.col.resize(size_t(ptr[n]));
    A.val.resize(size_t(ptr[n]));
#pragma omp parallel for schedule(static) if (ptr[n] > 200000)
    for (int64_t k = 0; k < ptr[n]; k++) {
        A.col[size_t(k)] = col[k];
        A.val[size_t(k)] = cd(val[2 * k], val[2 * k + 1]);
    }
    return build_impl<cd>(c, std::move(A), o, tune().theta_cplx, tune().coarse_cplx);
}
void mg_apply64(mk_ctx* c, Mg* g, const double* b, double* x, const int* skip) { vcycle<double>(c, g, 0, b, x, skip); }
void mg_apply32(mk_ctx* c, Mg* g, const float* b, float* x, const int* skip) { vcycle<float>(c, g, 0, b, x, skip); }
template <typename T>
void mg_apply_block(mk_ctx* c, Mg* g, int mw, const T* b, T* x, const int* skip) {
    MK_CHECK(!g->cplx, MK_ERR_UNSUPPORTED, "batched V-cycle: real hierarchies only");
    MK_CHECK(mw == 8 || mw == 16 || mw == 32 || mw == 64, MK_ERR_ARG, "batched V-cycle: block width");
    const int sh = mw == 8 ? 3 : mw == 16 ? 4 : mw == 32 ? 5 : 6;
    for (size_t l = 0; l < g->lv.size(); l++) {
        MgLevel& L = g->lv[l];
        const size_t len = size_t(L.n) * mw;
        if constexpr (std::is_same<T, double>::value) {
            if (l > 0) { L.bb64.ensure(len); L.bx64.ensure(len); }
            if (l + 1 < g->lv.size()) { L.br64.ensure(len); L.by64.ensure(len); }
        } else {
            if (l > 0) { L.bb32.ensure(len); L.bx32.ensure(len); }
            if (l + 1 < g->lv.size()) { L.br32.ensure(len); L.by32.ensure(len); }
        }
    }
    vcycle_block<T>(c, g, 0, mw, sh, b, x, skip);
}
template void mg_apply_block<float>(mk_ctx*, Mg*, int, const float*, float*, const int*);
template void mg_apply_block<double>(mk_ctx*, Mg*, int, const double*, double*, const int*);
int mg_levels(const Mg* g) { return int(g->lv.size()); }
uint64_t mg_uid(const Mg* g) { return g ? g->uid : 0; }
void mg_set_labels(Mg* g, const std::string& base) {
    for (size_t l = 0; l < g->lv.size(); l++) {
        const std::string s = base + "_mg" + std::to_string(l);
        if (g->lv[l].A) g->lv[l].A->label = s;
        if (g->lv[l].P) g->lv[l].P->label = s + "_P";
        if (g->lv[l].R) g->lv[l].R->label = s + "_R";
        if (g->lv[l].RR) g->lv[l].RR->label = s + "_RR";
    }
}
double mg_setup_ms(const Mg* g) { return g->setup_ms; }
void mg_set_setup_ms(Mg* g, double ms) { g->setup_ms = ms; }
int64_t mg_nnz(const Mg* g) { return g->nnz_total; }
int mg_level_rows(const Mg* g, int l) { return g->lv[size_t(l)].n; }

}  // namespace mk
