// metriko (MI355X build) -- PositiveDefiniteSolver on the GPU.
//
// Drop-in for the reference's core/solver/suitesparse/suitesparse_pdefinite.h:14-62
// (CHOLMOD simplicial LDL^T of a symmetric / Hermitian positive definite matrix): the
// constructor uploads the matrix, solve() runs multigrid-preconditioned PCG on the device to
// the runtime's tolerance (1e-10 FP64 relative residual by default).  A matrix that is not
// positive definite is reported when PCG breaks down (std::runtime_error, MK_ERR_NOT_SPD),
// i.e. at the first solve rather than in the constructor.
//
// Restated from komietty/metriko include/metriko/core/solver/suitesparse/suitesparse_pdefinite.h
// (the komietty/metriko authors), which is subject to the terms of the Mozilla Public License v.2.0.
#ifndef METRIKO_SUITESPARSE_PDEFINITE_H
#define METRIKO_SUITESPARSE_PDEFINITE_H
#include "suitesparse_utilities.h"

namespace metriko { namespace suitesparse {
    template <typename T>
    class PositiveDefiniteSolver : public LinearSolver<T> {
    public:
        explicit PositiveDefiniteSolver(const Sprs<T>& mat)
            : LinearSolver<T>(mat), opts(std::is_same<T, double>::value ? hip::runtime().igm_opts
                                                                          : hip::runtime().field_opts),
              dev(mat, opts) {}
        void solve(VecX<T>& x, const VecX<T>& rhs) override {
            x = VecX<T>::Zero(int(this->nRows));
            dev.solve(x, rhs, opts, &last);
        }
        mk_solve_opts opts;
        mk_stats last{};

    protected:
        detail::DeviceMatrix<T> dev;
    };
}}  // namespace metriko::suitesparse

#endif
