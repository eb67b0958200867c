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
#include "typedef.h"

namespace metriko {
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
        std::vector<int64_t> ptr{0};
        std::vector<int> col;
        std::vector<T> val;

        Csr() = default;
        Csr(int r, int c) : nr(r), nc(c), ptr(size_t(r) + 1, 0) {}
        int rows() const { return nr; }
        int cols() const { return nc; }
        int64_t nonZeros() const { return ptr.empty() ? 0 : ptr.back(); }
        void resize(int r, int c) { nr = r; nc = c; ptr.assign(size_t(r) + 1, 0); col.clear(); val.clear(); }

        // Eigen setFromTriplets semantics: duplicates are summed.
        void setFromTriplets(const std::vector<Triplet<T>>& T_) {
            std::vector<int64_t> cnt(size_t(nr) + 1, 0);
            for (auto& t : T_) cnt[size_t(t.r) + 1]++;
            for (int i = 0; i < nr; i++) cnt[i + 1] += cnt[i];
            std::vector<int> c(T_.size());
            std::vector<T> v(T_.size());
            std::vector<int64_t> pos(cnt.begin(), cnt.end() - 1);
            for (auto& t : T_) { int64_t p = pos[t.r]++; c[p] = t.c; v[p] = t.v; }
            ptr.assign(size_t(nr) + 1, 0);
            col.clear(); val.clear();
            col.reserve(T_.size()); val.reserve(T_.size());
            std::vector<int> idx;
            for (int i = 0; i < nr; i++) {
                idx.resize(size_t(cnt[i + 1] - cnt[i]));
                std::iota(idx.begin(), idx.end(), 0);
                const int64_t b = cnt[i];
                std::stable_sort(idx.begin(), idx.end(), [&](int a, int bb) { return c[b + a] < c[b + bb]; });
                int last = -1;
                for (int k : idx) {
                    if (c[b + k] == last) val.back() += v[b + k];
                    else { col.push_back(c[b + k]); val.push_back(v[b + k]); last = c[b + k]; }
                }
                ptr[i + 1] = int64_t(col.size());
            }
        }

        // Drops |x| <= ref*eps entries (Eigen prune(ref, eps) with eps = 1e-12).
        void prune(double ref = 0.0, double eps = 1e-12) {
            const double th = std::abs(ref) * eps;
            int64_t w = 0;
            std::vector<int64_t> np(size_t(nr) + 1, 0);
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
            Csr t(nc, nr);
            std::vector<int64_t> cnt(size_t(nc) + 1, 0);
            for (int64_t k = 0; k < nonZeros(); k++) cnt[size_t(col[k]) + 1]++;
            for (int i = 0; i < nc; i++) cnt[i + 1] += cnt[i];
            t.ptr = cnt;
            t.col.resize(col.size());
            t.val.resize(val.size());
            std::vector<int64_t> pos(cnt.begin(), cnt.end() - 1);
            for (int i = 0; i < nr; i++)
                for (int64_t k = ptr[i]; k < ptr[i + 1]; k++) {
                    int64_t p = pos[col[k]]++;
                    t.col[p] = i;
                    t.val[p] = val[k];
                }
            return t;
        }

        void multiply(const T* x, T* y) const {
            for (int i = 0; i < nr; i++) {
                T s = T();
                for (int64_t k = ptr[i]; k < ptr[i + 1]; k++) s += val[k] * x[col[k]];
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

    // C = A * B (Gustavson with a dense accumulator; sorted output columns)
    template<typename T>
    Csr<T> spgemm(const Csr<T>& A, const Csr<T>& B) {
        if (A.nc != B.nr) throw std::invalid_argument("spgemm: shape mismatch");
        Csr<T> C(A.nr, B.nc);
        std::vector<T> acc(size_t(B.nc), T());
        std::vector<int> mark(size_t(B.nc), -1);
        std::vector<int> cols;
        for (int i = 0; i < A.nr; i++) {
            cols.clear();
            for (int64_t ka = A.ptr[i]; ka < A.ptr[i + 1]; ka++) {
                const int j = A.col[ka];
                const T a = A.val[ka];
                for (int64_t kb = B.ptr[j]; kb < B.ptr[j + 1]; kb++) {
                    const int c = B.col[kb];
                    if (mark[c] != i) { mark[c] = i; acc[c] = a * B.val[kb]; cols.push_back(c); }
                    else acc[c] += a * B.val[kb];
                }
            }
            std::sort(cols.begin(), cols.end());
            for (int c : cols) { C.col.push_back(c); C.val.push_back(acc[c]); }
            C.ptr[i + 1] = int64_t(C.col.size());
        }
        return C;
    }

    template<typename T>
    Csr<T> operator*(const Csr<T>& A, const Csr<T>& B) { return spgemm(A, B); }

    template<typename T>
    Csr<T> operator+(const Csr<T>& A, const Csr<T>& B) {
        std::vector<Triplet<T>> t;
        t.reserve(size_t(A.nonZeros() + B.nonZeros()));
        for (int i = 0; i < A.nr; i++) for (int64_t k = A.ptr[i]; k < A.ptr[i + 1]; k++) t.emplace_back(i, A.col[k], A.val[k]);
        for (int i = 0; i < B.nr; i++) for (int64_t k = B.ptr[i]; k < B.ptr[i + 1]; k++) t.emplace_back(i, B.col[k], B.val[k]);
        Csr<T> C(A.nr, A.nc);
        C.setFromTriplets(t);
        return C;
    }

    using SprsD = Csr<double>;
    using SprsC = Csr<complex>;
    using TripD = Triplet<double>;
    using TripC = Triplet<complex>;
    using TripI = Triplet<int>;
}

#endif
