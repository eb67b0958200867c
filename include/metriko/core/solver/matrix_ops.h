// metriko (MI355X build) -- small dense / sparse helpers of the IGM setup.
//
// Drop-in for the reference's core/solver/matrix_ops.h plus the Eigen FullPivLU
// kernel that iter_rounding_init.h:129 calls.
#ifndef METRIKO_SOLVER_COMMON_H
#define METRIKO_SOLVER_COMMON_H
#include <algorithm>
#include <cmath>
#include <limits>
#include "metriko/core/common/sparse.h"

namespace metriko {
    // Deterministic maximal linearly independent row subset, kept in row order: row i
    // stays when its residual against the kept rows (two Gram-Schmidt passes) exceeds
    // eps*1e3*ncols*||row||.  Stands in for reduce_to_linearly_independent
    // (matrix_ops.h:12-28), whose row choice follows Eigen SparseQR<COLAMD>'s column
    // permutation; that ordering is not restated (the kept subset spans the same row
    // space, which is all the downstream kernel basis depends on up to its pivoting).
    inline void reduce_to_linearly_independent(SprsD& mat) {
        if (mat.rows() == 0) return;
        std::vector<int> cols;
        for (int64_t k = 0; k < mat.nonZeros(); k++) cols.push_back(mat.col[k]);
        std::sort(cols.begin(), cols.end());
        cols.erase(std::unique(cols.begin(), cols.end()), cols.end());
        const int nc = int(cols.size());
        std::vector<int> cmap(size_t(mat.cols()), -1);
        for (int j = 0; j < nc; j++) cmap[cols[j]] = j;
        const double tol = std::numeric_limits<double>::epsilon() * 1e3 * std::max(nc, 1);
        std::vector<std::vector<double>> Q;
        std::vector<int> keep;
        std::vector<double> r(static_cast<size_t>(nc));
        for (int i = 0; i < mat.rows(); i++) {
            std::fill(r.begin(), r.end(), 0.0);
            for (int64_t k = mat.ptr[i]; k < mat.ptr[i + 1]; k++) r[cmap[mat.col[k]]] = mat.val[k];
            double nrm = 0;
            for (double x : r) nrm += x * x;
            nrm = std::sqrt(nrm);
            if (nrm == 0.0) continue;
            for (int pass = 0; pass < 2; pass++)
                for (auto& q : Q) {
                    double d = 0;
                    for (int j = 0; j < nc; j++) d += q[j] * r[j];
                    for (int j = 0; j < nc; j++) r[j] -= d * q[j];
                }
            double rn = 0;
            for (double x : r) rn += x * x;
            rn = std::sqrt(rn);
            if (rn > tol * nrm) {
                for (double& x : r) x /= rn;
                Q.push_back(r);
                keep.push_back(i);
            }
        }
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
