// metriko (MI355X build) -- compressed-sparse-row matrices for the host side.
//
// Stands in for Eigen::SparseMatrix in the drop-in headers: triplet assembly
// that sums duplicates like setFromTriplets, prune of exact zeros, transpose,
// sparse products (Gustavson, row by row) and row selection.  Column indices
// inside a row are kept sorted.
#ifndef METRIKO_SPARSE_H
#define METRIKO_SPARSE_H
#include <algorithm>
#include <numeric>
#include <stdexcept>
#ifdef _OPENMP
#include <omp.h>
#endif
#include "hvec.h"
#include "typedef.h"

namespace metriko {
    // Product of two values; complex ones by the textbook formula, inline.  For finite operands
    // this is the value of the compiler runtime's complex multiply, which adds only the
    // NaN/infinity recovery and costs a call per product.
    template<typename T>
    inline T mulv(const T& a, const T& b) { return a * b; }
    inline std::complex<double> mulv(const std::complex<double>& a, const std::complex<double>& b) {
        const double ac = a.real() * b.real(), bd = a.imag() * b.imag();
        const double ad = a.real() * b.imag(), bc = a.imag() * b.real();
        return {ac - bd, ad + bc};
    }

    template<typename T>
    struct Triplet {
        int r, c;
        T v;
        Triplet(int r_, int c_, T v_) : r(r_), c(c_), v(v_) {}
        int row() const { return r; }
        int col() const { return c; }
        T value() const { return v; }
    };

    template<typename T>
    struct Csr {
        int nr = 0, nc = 0;
        HVec<int64_t> ptr{0};
        HVec<int> col;
        HVec<T> val;

        Csr() = default;
        Csr(int r, int c) : nr(r), nc(c), ptr(size_t(r) + 1, 0) {}
        int rows() const { return nr; }
        int cols() const { return nc; }
        int64_t nonZeros() const { return ptr.empty() ? 0 : ptr.back(); }
        void resize(int r, int c) { nr = r; nc = c; ptr.assign(size_t(r) + 1, 0); col.clear(); val.clear(); }

        // Eigen setFromTriplets semantics: duplicates are summed (in input order within a row).
        void setFromTriplets(const std::vector<Triplet<T>>& T_) {
            std::vector<int64_t> cnt(size_t(nr) + 1, 0);
            for (auto& t : T_) cnt[size_t(t.r) + 1]++;
            for (int i = 0; i < nr; i++) cnt[i + 1] += cnt[i];
            std::vector<int> c(T_.size());
            std::vector<T> v(T_.size());
            std::vector<int64_t> pos(cnt.begin(), cnt.end() - 1);
            for (auto& t : T_) { int64_t p = pos[t.r]++; c[p] = t.c; v[p] = t.v; }
            // rows are independent: sort + merge each (two passes: sizes, then values)
            HVec<int64_t> uq(size_t(nr) + 1, 0);
#pragma omp parallel if (T_.size() > 100000)
            {
                std::vector<int> idx, cs;
                std::vector<T> vs;
#pragma omp for schedule(dynamic, 512)
                for (int i = 0; i < nr; i++) {
                    const int64_t b = cnt[i], e = cnt[i + 1];
                    if (e - b <= 1) { uq[i + 1] = e - b; continue; }
                    idx.resize(size_t(e - b));
                    std::iota(idx.begin(), idx.end(), 0);
                    std::stable_sort(idx.begin(), idx.end(), [&](int a, int bb) { return c[b + a] < c[b + bb]; });
                    // reorder the row in place (sorted, stable) and count distinct columns
                    cs.resize(idx.size());
                    vs.resize(idx.size());
                    for (size_t k = 0; k < idx.size(); k++) { cs[k] = c[b + idx[k]]; vs[k] = v[b + idx[k]]; }
                    int64_t u = 0;
                    for (size_t k = 0; k < cs.size(); k++) {
                        c[b + int64_t(k)] = cs[k];
                        v[b + int64_t(k)] = vs[k];
                        if (k == 0 || cs[k] != cs[k - 1]) u++;
                    }
                    uq[i + 1] = u;
                }
            }
            for (int i = 0; i < nr; i++) uq[i + 1] += uq[i];
            ptr = uq;
            col.assign(size_t(uq[nr]), 0);
            val.assign(size_t(uq[nr]), T());
#pragma omp parallel for schedule(dynamic, 512) if (T_.size() > 100000)
            for (int i = 0; i < nr; i++) {
                int64_t w = ptr[i] - 1;
                int last = -1;
                for (int64_t k = cnt[i]; k < cnt[i + 1]; k++) {
                    if (c[k] == last) val[w] += v[k];
                    else { ++w; col[w] = c[k]; val[w] = v[k]; last = c[k]; }
                }
            }
        }

        // Drops |x| <= ref*eps entries (Eigen prune(ref, eps) with eps = 1e-12).
        void prune(double ref = 0.0, double eps = 1e-12) {
            const double th = std::abs(ref) * eps;
            int64_t w = 0;
            HVec<int64_t> np(size_t(nr) + 1, 0);
            for (int i = 0; i < nr; i++) {
                for (int64_t k = ptr[i]; k < ptr[i + 1]; k++) {
                    if (std::abs(val[k]) > th) { col[w] = col[k]; val[w] = val[k]; w++; }
                }
                np[i + 1] = w;
            }
            col.resize(w); val.resize(w); ptr = np;
        }
        Csr pruned() const { Csr c = *this; c.prune(); return c; }

        Csr transpose() const {
            // offsets fit 32 bits for all but huge matrices: half the count arrays' traffic
            if (nonZeros() < (int64_t(1) << 31)) return transpose_impl<int>();
            return transpose_impl<int64_t>();
        }
        template<typename I>
        Csr transpose_impl() const {
            Csr t(nc, nr);
            const int64_t nz = nonZeros();
            t.col.resize(size_t(nz));
            t.val.resize(size_t(nz));
            // rows split into chunks: per-chunk column counts give every entry its slot, so the
            // result (rows ascending inside each column) does not depend on the chunking
            int nch = 1;
#ifdef _OPENMP
            nch = std::max(1, std::min(omp_get_max_threads(), int(nz / 200000) + 1));
#endif
            std::vector<std::vector<I>> cc(static_cast<size_t>(nch));
            const int step = (nr + nch - 1) / std::max(nch, 1);
#pragma omp parallel for schedule(static, 1)
            for (int q = 0; q < nch; q++) {
                cc[size_t(q)].assign(size_t(nc), I(0));
                const int r1 = std::min(nr, (q + 1) * step);
                for (int i = q * step; i < r1; i++)
                    for (int64_t k = ptr[i]; k < ptr[i + 1]; k++) cc[size_t(q)][size_t(col[k])]++;
            }
            t.ptr.resize(size_t(nc) + 1);
            t.ptr[0] = 0;
            // offsets: column j's entries go chunk by chunk; the columns are split into nch
            // ranges whose totals are scanned first, so the per-column pass runs in parallel
            const int cstep = (nc + nch - 1) / std::max(nch, 1);
            std::vector<int64_t> rtot(size_t(nch) + 1, 0);
#pragma omp parallel for schedule(static, 1)
            for (int q = 0; q < nch; q++) {
                int64_t s = 0;
                const int c1 = std::min(nc, (q + 1) * cstep);
                for (int j = q * cstep; j < c1; j++)
                    for (int p = 0; p < nch; p++) s += cc[size_t(p)][size_t(j)];
                rtot[size_t(q) + 1] = s;
            }
            for (int q = 0; q < nch; q++) rtot[size_t(q) + 1] += rtot[size_t(q)];
#pragma omp parallel for schedule(static, 1)
            for (int q = 0; q < nch; q++) {
                int64_t s = rtot[size_t(q)];
                const int c1 = std::min(nc, (q + 1) * cstep);
                for (int j = q * cstep; j < c1; j++) {
                    for (int p = 0; p < nch; p++) { const int64_t x = cc[size_t(p)][size_t(j)]; cc[size_t(p)][size_t(j)] = I(s); s += x; }
                    t.ptr[size_t(j) + 1] = s;
                }
            }
#pragma omp parallel for schedule(static, 1)
            for (int q = 0; q < nch; q++) {
                std::vector<I>& pos = cc[size_t(q)];
                const int r1 = std::min(nr, (q + 1) * step);
                for (int i = q * step; i < r1; i++)
                    for (int64_t k = ptr[i]; k < ptr[i + 1]; k++) {
                        const int64_t p = pos[size_t(col[k])]++;
                        t.col[p] = i;
                        t.val[p] = val[k];
                    }
            }
            return t;
        }

        void multiply(const T* x, T* y) const {
#pragma omp parallel for schedule(static, 4096) if (nonZeros() > 200000)
            for (int i = 0; i < nr; i++) {
                T s = T();
                for (int64_t k = ptr[i]; k < ptr[i + 1]; k++) s += mulv(val[k], x[col[k]]);
                y[i] = s;
            }
        }
        VecX<T> operator*(const VecX<T>& x) const { VecX<T> y(nr); multiply(x.data(), y.data()); return y; }

        Csr scaled(T s) const { Csr c = *this; for (auto& v : c.val) v *= s; return c; }

        // rows idx (in the given order) as a new matrix (igl::slice(A, idx, 1, B))
        Csr rows_of(const std::vector<int>& idx) const {
            Csr r(int(idx.size()), nc);
            for (size_t i = 0; i < idx.size(); i++) {
                for (int64_t k = ptr[idx[i]]; k < ptr[idx[i] + 1]; k++) { r.col.push_back(col[k]); r.val.push_back(val[k]); }
                r.ptr[i + 1] = int64_t(r.col.size());
            }
            return r;
        }

        static Csr identity(int n, T s = T(1)) {
            Csr I(n, n);
            I.col.resize(n); I.val.assign(n, s);
            for (int i = 0; i < n; i++) { I.col[i] = i; I.ptr[i + 1] = i + 1; }
            return I;
        }
    };

    // C = A * B (Gustavson; sorted output columns).  Rows are independent: a symbolic pass
    // sizes them, a numeric pass fills them.  Short rows (few products) accumulate by a
    // stable sort of their (column, product) list, long rows in a dense accumulator; both
    // add the products of a column in the same order (A's row order, then B's), so the
    // result does not depend on the path or on the thread count.
    template<typename T>
    Csr<T> spgemm(const Csr<T>& A, const Csr<T>& B) {
        if (A.nc != B.nr) throw std::invalid_argument("spgemm: shape mismatch");
        Csr<T> C(A.nr, B.nc);
        const bool par = A.nonZeros() > 100000;
        constexpr int64_t kShort = 32;  // products per row up to which the sort path is used
        auto work = [&](int i) {
            int64_t w = 0;
            for (int64_t ka = A.ptr[i]; ka < A.ptr[i + 1]; ka++) w += B.ptr[A.col[ka] + 1] - B.ptr[A.col[ka]];
            return w;
        };
        HVec<int64_t> cnt(size_t(A.nr) + 1, 0);
#pragma omp parallel if (par)
        {
            std::vector<int> mark, cols;
#pragma omp for schedule(dynamic, 1024)
            for (int i = 0; i < A.nr; i++) {
                int64_t n = 0;
                if (work(i) <= kShort) {
                    cols.clear();
                    for (int64_t ka = A.ptr[i]; ka < A.ptr[i + 1]; ka++) {
                        const int j = A.col[ka];
                        for (int64_t kb = B.ptr[j]; kb < B.ptr[j + 1]; kb++) cols.push_back(B.col[kb]);
                    }
                    for (size_t q = 1; q < cols.size(); q++)  // insertion sort (a few entries)
                        for (size_t r = q; r > 0 && cols[r - 1] > cols[r]; r--) std::swap(cols[r - 1], cols[r]);
                    n = int64_t(std::unique(cols.begin(), cols.end()) - cols.begin());
                } else {
                    if (mark.empty()) mark.assign(size_t(B.nc), -1);
                    for (int64_t ka = A.ptr[i]; ka < A.ptr[i + 1]; ka++) {
                        const int j = A.col[ka];
                        for (int64_t kb = B.ptr[j]; kb < B.ptr[j + 1]; kb++)
                            if (mark[B.col[kb]] != i) { mark[B.col[kb]] = i; n++; }
                    }
                }
                cnt[size_t(i) + 1] = n;
            }
        }
        for (int i = 0; i < A.nr; i++) cnt[size_t(i) + 1] += cnt[size_t(i)];
        C.ptr = std::move(cnt);
        C.col.resize(size_t(C.ptr[size_t(A.nr)]));
        C.val.resize(size_t(C.ptr[size_t(A.nr)]));
#pragma omp parallel if (par)
        {
            std::vector<T> acc;
            std::vector<int> mark;
            std::vector<int> cols;
            std::vector<std::pair<int, T>> pr;
#pragma omp for schedule(dynamic, 1024)
            for (int i = 0; i < A.nr; i++) {
                int64_t w = C.ptr[size_t(i)];
                if (work(i) <= kShort) {
                    pr.clear();
                    for (int64_t ka = A.ptr[i]; ka < A.ptr[i + 1]; ka++) {
                        const int j = A.col[ka];
                        const T a = A.val[ka];
                        for (int64_t kb = B.ptr[j]; kb < B.ptr[j + 1]; kb++) pr.emplace_back(B.col[kb], mulv(a, B.val[kb]));
                    }
                    for (size_t q = 1; q < pr.size(); q++)  // stable insertion sort by column
                        for (size_t r = q; r > 0 && pr[r - 1].first > pr[r].first; r--) std::swap(pr[r - 1], pr[r]);
                    for (size_t q = 0; q < pr.size(); q++) {
                        if (q > 0 && pr[q].first == pr[q - 1].first) C.val[size_t(w - 1)] += pr[q].second;
                        else { C.col[size_t(w)] = pr[q].first; C.val[size_t(w)] = pr[q].second; w++; }
                    }
                    continue;
                }
                if (acc.empty()) { acc.assign(size_t(B.nc), T()); mark.assign(size_t(B.nc), -1); }
                cols.clear();
                for (int64_t ka = A.ptr[i]; ka < A.ptr[i + 1]; ka++) {
                    const int j = A.col[ka];
                    const T a = A.val[ka];
                    for (int64_t kb = B.ptr[j]; kb < B.ptr[j + 1]; kb++) {
                        const int c = B.col[kb];
                        if (mark[c] != i) { mark[c] = i; acc[c] = mulv(a, B.val[kb]); cols.push_back(c); }
                        else acc[c] += mulv(a, B.val[kb]);
                    }
                }
                std::sort(cols.begin(), cols.end());
                for (int c : cols) { C.col[size_t(w)] = c; C.val[size_t(w)] = acc[c]; w++; }
            }
        }
        return C;
    }

    template<typename T>
    Csr<T> operator*(const Csr<T>& A, const Csr<T>& B) { return spgemm(A, B); }

    // C = a A + b B by merging the sorted rows (entries present in either operand are kept,
    // zeros included, like a triplet assembly of both).  Rows are merged in parallel.
    template<typename T>
    Csr<T> add(const Csr<T>& A, T a, const Csr<T>& B, T b) {
        if (A.nr != B.nr || A.nc != B.nc) throw std::invalid_argument("sparse add: shape mismatch");
        Csr<T> C(A.nr, A.nc);
        const bool par = A.nonZeros() + B.nonZeros() > 200000;
        HVec<int64_t> cnt(size_t(A.nr) + 1, 0);
#pragma omp parallel for schedule(static, 4096) if (par)
        for (int i = 0; i < A.nr; i++) {
            int64_t p = A.ptr[i], q = B.ptr[i], n = 0;
            while (p < A.ptr[i + 1] || q < B.ptr[i + 1]) {
                if (q >= B.ptr[i + 1] || (p < A.ptr[i + 1] && A.col[p] < B.col[q])) p++;
                else if (p >= A.ptr[i + 1] || B.col[q] < A.col[p]) q++;
                else { p++; q++; }
                n++;
            }
            cnt[size_t(i) + 1] = n;
        }
        for (int i = 0; i < A.nr; i++) cnt[size_t(i) + 1] += cnt[size_t(i)];
        C.ptr = std::move(cnt);
        C.col.resize(size_t(C.ptr[size_t(A.nr)]));
        C.val.resize(size_t(C.ptr[size_t(A.nr)]));
#pragma omp parallel for schedule(static, 4096) if (par)
        for (int i = 0; i < A.nr; i++) {
            int64_t p = A.ptr[i], q = B.ptr[i], w = C.ptr[i];
            while (p < A.ptr[i + 1] || q < B.ptr[i + 1]) {
                if (q >= B.ptr[i + 1] || (p < A.ptr[i + 1] && A.col[p] < B.col[q])) {
                    C.col[w] = A.col[p]; C.val[w] = a * A.val[p]; p++;
                } else if (p >= A.ptr[i + 1] || B.col[q] < A.col[p]) {
                    C.col[w] = B.col[q]; C.val[w] = b * B.val[q]; q++;
                } else {
                    C.col[w] = A.col[p]; C.val[w] = a * A.val[p] + b * B.val[q]; p++; q++;
                }
                w++;
            }
        }
        return C;
    }

    template<typename T>
    Csr<T> operator+(const Csr<T>& A, const Csr<T>& B) { return add(A, T(1), B, T(1)); }

    using SprsD = Csr<double>;
    using SprsC = Csr<complex>;
    using TripD = Triplet<double>;
    using TripC = Triplet<complex>;
    using TripI = Triplet<int>;
}

#endif
