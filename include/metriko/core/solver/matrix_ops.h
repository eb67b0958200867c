// metriko (MI355X build) -- small dense / sparse helpers of the IGM setup.
//
// Drop-in for the reference's core/solver/matrix_ops.h plus the Eigen FullPivLU
// kernel that iter_rounding_init.h:129 calls.
//
// Restated from komietty/metriko include/metriko/core/solver/matrix_ops.h
// (Copyright (C) 2021 Amir Vaxman, Copyright (C) 2025 Saki Komikado), which is subject to
// the terms of the Mozilla Public License v.2.0.
#ifndef METRIKO_SOLVER_COMMON_H
#define METRIKO_SOLVER_COMMON_H
#include <algorithm>
#include <cmath>
#include <limits>
#include "metriko/core/common/sparse.h"
#include "metriko/core/solver/colamd.h"

namespace metriko {
    // reduce_to_linearly_independent (matrix_ops.h:12-28): the reference factorises C^T with
    // Eigen::SparseQR<SprsD, COLAMDOrdering<int>> and keeps the rows of C named by the first
    // rank entries of colsPermutation(), in that order.  SparseQR visits the columns of C^T
    // (the rows of C) in COLAMD's order; a left-looking Householder step whose norm |beta| --
    // the distance of the column to the span of the columns accepted before it -- is below
    // 20 (m + n) max_j ||C^T(:,j)|| eps moves the column behind the accepted ones.  So the kept
    // rows are the greedy independent subset of the COLAMD order, in that order.  Here the
    // distances are computed per connected component of the constraint graph (rows sharing a
    // column; disjoint supports cannot be dependent), by Householder steps on the component's
    // dense block, with the global threshold.  Explicit zeros (from the structural product that
    // built C) stay in the pattern: COLAMD sees them and so do the kept rows.
    inline void reduce_to_linearly_independent(SprsD& mat) {
        if (mat.rows() == 0) return;
        const int nr = mat.rows(), nc = mat.cols();
        // C^T in compressed-column form is C in compressed-row form
        const std::vector<int> p = colamd::order(nc, nr, mat.ptr, mat.col);
        double max2 = 0.0;
        for (int i = 0; i < nr; i++) {
            double s2 = 0.0;
            for (int64_t k = mat.ptr[i]; k < mat.ptr[i + 1]; k++) s2 += mat.val[k] * mat.val[k];
            max2 = std::max(max2, std::sqrt(s2));
        }
        if (max2 == 0.0) max2 = 1.0;
        const double thr = 20.0 * (double(nc) + double(nr)) * max2 * std::numeric_limits<double>::epsilon();
        // components: union-find over the columns each row touches
        std::vector<int> uf(static_cast<size_t>(nc));
        for (int j = 0; j < nc; j++) uf[j] = j;
        auto find = [&](int x) { while (uf[x] != x) { uf[x] = uf[uf[x]]; x = uf[x]; } return x; };
        for (int i = 0; i < nr; i++)
            for (int64_t k = mat.ptr[i] + 1; k < mat.ptr[i + 1]; k++) {
                const int a = find(mat.col[mat.ptr[i]]), b = find(mat.col[k]);
                if (a != b) uf[std::max(a, b)] = std::min(a, b);
            }
        // per component: its rows in COLAMD order, its columns ascending
        std::vector<int> comp_of_row(static_cast<size_t>(nr), -1), comp_id(static_cast<size_t>(nc), -1);
        int ncomp = 0;
        std::vector<std::vector<int>> crows, ccols;
        for (int t = 0; t < nr; t++) {
            const int i = p[t];
            if (mat.ptr[i] == mat.ptr[i + 1]) continue;  // an empty row is never kept
            const int root = find(mat.col[mat.ptr[i]]);
            if (comp_id[root] < 0) { comp_id[root] = ncomp++; crows.emplace_back(); ccols.emplace_back(); }
            comp_of_row[i] = comp_id[root];
            crows[comp_id[root]].push_back(i);
        }
        for (int j = 0; j < nc; j++) {
            const int r = find(j);
            if (comp_id[r] >= 0) ccols[comp_id[r]].push_back(j);
        }
        std::vector<char> accepted(static_cast<size_t>(nr), 0);
        std::vector<int> local(static_cast<size_t>(nc), -1);
        for (int g = 0; g < ncomp; g++) {
            const std::vector<int>& cols = ccols[g];
            for (size_t a = 0; a < cols.size(); a++) local[cols[a]] = int(a);
            const int m = int(cols.size());
            std::vector<std::vector<double>> H;  // accepted Householder vectors (unit entry at its step)
            std::vector<double> tau;
            std::vector<double> v(static_cast<size_t>(m));
            for (int i : crows[g]) {
                std::fill(v.begin(), v.end(), 0.0);
                for (int64_t k = mat.ptr[i]; k < mat.ptr[i + 1]; k++) v[local[mat.col[k]]] = mat.val[k];
                for (size_t h = 0; h < H.size(); h++) {
                    double d = 0.0;
                    for (int a = 0; a < m; a++) d += H[h][a] * v[a];
                    d *= tau[h];
                    for (int a = 0; a < m; a++) v[a] -= d * H[h][a];
                }
                const int nz = int(H.size());
                if (nz >= m) continue;
                const double c0 = v[nz];
                double sq = 0.0;
                for (int a = nz + 1; a < m; a++) sq += v[a] * v[a];
                double beta, t;
                std::vector<double> hv(static_cast<size_t>(m), 0.0);
                hv[nz] = 1.0;
                if (sq == 0.0) {
                    beta = c0;
                    t = 0.0;
                } else {
                    beta = std::sqrt(c0 * c0 + sq);
                    if (c0 >= 0.0) beta = -beta;
                    for (int a = nz + 1; a < m; a++) hv[a] = v[a] / (c0 - beta);
                    t = (beta - c0) / beta;
                }
                if (std::abs(beta) >= thr) {
                    H.push_back(std::move(hv));
                    tau.push_back(t);
                    accepted[i] = 1;
                }
            }
            for (int j : cols) local[j] = -1;
        }
        std::vector<int> keep;
        for (int t = 0; t < nr; t++)
            if (accepted[p[t]]) keep.push_back(p[t]);
        mat = mat.rows_of(keep);
    }

    // Eigen 3.4 FullPivLU<MatrixXd>(A).kernel() (called at iter_rounding_init.h:129):
    // complete pivoting (largest |entry| of the trailing block, first in column-major
    // order on ties), rank = #pivots above maxPivot * eps * min(rows, cols), kernel by
    // back substitution on the leading triangle, rows placed by the column permutation.
    // A is rows x cols, row-major; returns cols x max(1, dimker), row-major.
    inline MatXd fullpiv_lu_kernel(const MatXd& A, int* rank_out = nullptr) {
        const int rows = A.rows(), cols = A.cols(), size = std::min(rows, cols);
        MatXd lu = A;
        std::vector<int> q(static_cast<size_t>(cols));
        for (int j = 0; j < cols; j++) q[j] = j;
        int nonzero = size;
        double maxpiv = 0.0;
        for (int k = 0; k < size; k++) {
            double big = -1.0;
            int rb = k, cb = k;
            for (int j = k; j < cols; j++)
                for (int i = k; i < rows; i++) {
                    const double a = std::abs(lu(i, j));
                    if (a > big) { big = a; rb = i; cb = j; }
                }
            if (big == 0.0) { nonzero = k; break; }
            if (big > maxpiv) maxpiv = big;
            if (rb != k) for (int j = 0; j < cols; j++) std::swap(lu(k, j), lu(rb, j));
            if (cb != k) { for (int i = 0; i < rows; i++) std::swap(lu(i, k), lu(i, cb)); std::swap(q[k], q[cb]); }
            const double piv = lu(k, k);
            for (int i = k + 1; i < rows; i++) lu(i, k) /= piv;
            for (int i = k + 1; i < rows; i++) {
                const double l = lu(i, k);
                if (l == 0.0) continue;
                for (int j = k + 1; j < cols; j++) lu(i, j) -= l * lu(k, j);
            }
        }
        const double thr = maxpiv * std::numeric_limits<double>::epsilon() * size;
        std::vector<int> piv;
        for (int i = 0; i < nonzero; i++) if (std::abs(lu(i, i)) > thr) piv.push_back(i);
        const int rank = int(piv.size());
        if (rank_out) *rank_out = rank;
        const int dimker = cols - rank;
        if (dimker == 0) return MatXd::Zero(cols, 1);
        MatXd m(rank, cols);
        for (int i = 0; i < rank; i++)
            for (int j = i; j < cols; j++) m(i, j) = lu(piv[i], j);
        for (int i = 0; i < rank; i++)
            if (piv[i] != i) for (int r = 0; r < rank; r++) std::swap(m(r, i), m(r, piv[i]));
        // back substitution: m[:, :rank] upper triangular, solve for the trailing columns
        for (int c = rank; c < cols; c++)
            for (int i = rank - 1; i >= 0; i--) {
                double s = m(i, c);
                for (int j = i + 1; j < rank; j++) s -= m(i, j) * m(j, c);
                m(i, c) = s / m(i, i);
            }
        for (int i = rank - 1; i >= 0; i--)
            if (piv[i] != i) for (int r = 0; r < rank; r++) std::swap(m(r, i), m(r, piv[i]));
        MatXd K = MatXd::Zero(cols, dimker);
        for (int i = 0; i < rank; i++)
            for (int k = 0; k < dimker; k++) K(q[i], k) = -m(i, rank + k);
        for (int k = 0; k < dimker; k++) K(q[rank + k], k) = 1.0;
        return K;
    }

    // sparse_block (matrix_ops.h:30-57): vertical / block concatenation
    template<typename Scalar>
    void sparse_block(const MatXi& idcs, const std::vector<Csr<Scalar>*>& mats, Csr<Scalar>& result) {
        const int ro = idcs.rows(), co = idcs.cols();
        std::vector<int> roff(size_t(ro), 0), coff(size_t(co), 0);
        for (int i = 1; i < ro; i++) roff[i] = roff[i - 1] + mats[idcs(i - 1, 0)]->rows();
        for (int i = 1; i < co; i++) coff[i] = coff[i - 1] + mats[idcs(0, i - 1)]->cols();
        result.resize(roff[ro - 1] + mats[idcs(ro - 1, 0)]->rows(), coff[co - 1] + mats[idcs(0, co - 1)]->cols());
        std::vector<Triplet<Scalar>> T;
        for (int i = 0; i < ro; i++)
            for (int j = 0; j < co; j++) {
                const Csr<Scalar>& M = *mats[i];
                for (int r = 0; r < M.rows(); r++)
                    for (int64_t k = M.ptr[r]; k < M.ptr[r + 1]; k++) T.emplace_back(roff[i] + r, coff[j] + M.col[k], M.val[k]);
            }
        result.setFromTriplets(T);
    }
}

#endif
