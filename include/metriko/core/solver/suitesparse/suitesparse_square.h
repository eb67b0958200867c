// metriko (MI355X build) -- SquareSolver on the GPU.
//
// Drop-in for the reference's core/solver/suitesparse/suitesparse_square.h:15-49 (UMFPACK LU
// of a general square matrix).  On the device: a Hermitian matrix is solved directly by PCG
// (its callers in the reference pass Hermitian positive definite systems); any other
// nonsingular matrix through the normal equations A^H A x = A^H b, with A^H A formed once in
// the constructor (host sparse product) and solved by multigrid PCG, A^H b by a device SpMV.
// The normal equations square the condition number: the residual of A x = b is that of the
// normal equations times up to cond(A).
//
// Restated from komietty/metriko include/metriko/core/solver/suitesparse/suitesparse_square.h
// (the komietty/metriko authors), which is subject to the terms of the Mozilla Public License v.2.0.
#ifndef METRIKO_SUITESPARSE_SQUARE_H
#define METRIKO_SUITESPARSE_SQUARE_H
#include "suitesparse_utilities.h"

namespace metriko { namespace suitesparse {
    template <typename T>
    class SquareSolver : public LinearSolver<T> {
    public:
        explicit SquareSolver(const Sprs<T>& mat)
            : LinearSolver<T>(mat), opts(std::is_same<T, double>::value ? hip::runtime().igm_opts
                                                                          : hip::runtime().field_opts) {
            if (mat.rows() != mat.cols()) throw std::runtime_error("metriko: SquareSolver needs a square matrix");
            hermitian = detail::is_hermitian(mat);
            if (hermitian) {
                dev = detail::DeviceMatrix<T>(mat, opts);
            } else {
                const Sprs<T> AH = detail::adjoint(mat);
                adj = detail::DeviceMatrix<T>(AH, opts);
                dev = detail::DeviceMatrix<T>(spgemm(AH, mat), opts);
            }
        }
        void solve(VecX<T>& x, const VecX<T>& rhs) override {
            x = VecX<T>::Zero(int(this->nRows));
            if (hermitian) dev.solve(x, rhs, opts, &last);
            else dev.solve(x, adj.apply(rhs), opts, &last);
        }
        bool hermitian = false;
        mk_solve_opts opts;
        mk_stats last{};

    protected:
        detail::DeviceMatrix<T> dev, adj;
    };
}}  // namespace metriko::suitesparse

#endif
