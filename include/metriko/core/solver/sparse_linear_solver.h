// metriko (MI355X build) -- solver seam of the field stage.
//
// Drop-in for the reference's core/solver/sparse_linear_solver.h.  Where the
// reference factorises L once (CHOLMOD simplicial LDL^H or Eigen SparseLU,
// sparse_linear_solver.h:29-54) and back-substitutes nIter times, this build
// assembles L on the GPU and runs nIter warm-started PCG solves there
// (FP32 inner + FP64 refinement by default), with the normalisation and
// Rayleigh steps in FP64.
//
// Restated from komietty/metriko include/metriko/core/solver/sparse_linear_solver.h
// (Copyright (C) 2025 Saki Komikado), which is subject to the terms of the Mozilla Public License v.2.0.
#ifndef METRIKO_SPARSE_LINEAR_SOLVER_H
#define METRIKO_SPARSE_LINEAR_SOLVER_H
#include <cstdlib>
#include "metriko/core/common/sparse.h"
#include "metriko/core/solver/suitesparse/suitesparse_square.h"
#include "metriko/hip/runtime.h"

namespace metriko {
    // Eigen 3.4 VecXc::Random(n): every entry draws re then im from std::rand,
    // each mapped to -1 + 2*rand()/RAND_MAX (sparse_linear_solver.h:32,44).
    inline VecXc random_complex_vector(int n) {
        VecXc u(n);
        for (int i = 0; i < n; i++) {
            const double re = -1.0 + (2.0 * double(std::rand())) / double(RAND_MAX);
            const double im = -1.0 + (2.0 * double(std::rand())) / double(RAND_MAX);
            u[i] = complex(re, im);
        }
        return u;
    }

    // Device operator handle (RAII).
    struct DeviceOp {
        mk_op* op = nullptr;
        DeviceOp() = default;
        explicit DeviceOp(mk_op* o) : op(o) {}
        DeviceOp(const DeviceOp&) = delete;
        DeviceOp& operator=(const DeviceOp&) = delete;
        DeviceOp(DeviceOp&& o) noexcept : op(o.op) { o.op = nullptr; }
        DeviceOp& operator=(DeviceOp&& o) noexcept { if (op) mk_op_destroy(op); op = o.op; o.op = nullptr; return *this; }
        ~DeviceOp() { if (op) mk_op_destroy(op); }
    };

    inline DeviceOp upload_complex(const SprsC& A) {
        auto& rt = hip::runtime();
        mk_op* op = nullptr;
        hip::check(mk_op_create_complex(rt.get(), A.rows(), A.ptr.data(), A.col.data(),
                                        reinterpret_cast<const double*>(A.val.data()), &rt.field_opts, &op),
                   "mk_op_create_complex");
        return DeviceOp(op);
    }

    // x = argmin of the 50-step inverse power iteration on an operator already on the device.
    inline VecXc solveSmallestEig(const DeviceOp& L, const VecXd& mass, const VecXc& u0, int nIter = 50) {
        auto& rt = hip::runtime();
        VecXc x(int(mass.size()));
        hip::check(mk_field_smallest_eig(rt.get(), L.op, mass.data(), reinterpret_cast<const double*>(u0.data()), nIter,
                                         &rt.field_opts, reinterpret_cast<double*>(x.data()), &rt.last_field),
                   "mk_field_smallest_eig");
        return x;
    }

    // Reference signature (sparse_linear_solver.h:29): M is the (diagonal) mass matrix.
    inline VecXc solveSmallestEig(SprsC& L, SprsC& M, int nIter = 50) {
        VecXd mass(L.rows());
        for (int i = 0; i < M.rows(); i++)
            for (int64_t k = M.ptr[i]; k < M.ptr[i + 1]; k++)
                if (M.col[k] == i) mass[i] = M.val[k].real();
                else throw std::invalid_argument("solveSmallestEig: mass matrix must be diagonal");
        DeviceOp d = upload_complex(L);
        VecXc u = random_complex_vector(L.rows());
        return solveSmallestEig(d, mass, u, nIter);
    }

    // solveSquare (sparse_linear_solver.h:16-27): SquareSolver on the device -- PCG for the
    // Hermitian positive definite lhs the reference's callers pass (CurvatureAligned: L - 0*M),
    // the normal equations for any other nonsingular lhs.
    inline VecXc solveSquare(SprsC& lhs, VecXc& rhs) {
        suitesparse::SquareSolver<complex> solver(lhs);
        VecXc x;
        solver.solve(x, rhs);
        hip::runtime().last_field = solver.last;
        return x;
    }

    // Eigen 3.4 VectorXd::Random(n): -1 + 2*rand()/RAND_MAX per entry (std::rand).
    inline VecXd random_real_vector(int n) {
        VecXd u(n);
        for (int i = 0; i < n; i++) u[i] = -1.0 + (2.0 * double(std::rand())) / double(RAND_MAX);
        return u;
    }

    // Real solveSmallestEig (sparse_linear_solver.h:56-70): the reference factorises L once
    // (SimplicialLDLT) and runs nIter steps x = L^-1 (M u), u = x / sqrt(x^T M x), returning the
    // last x (not normalised).  Here L and M live on the device; each step is a PCG solve
    // (starting from the previous x) and two SpMVs.  The reference's per-step "iter i" trace on
    // stdout is not reproduced.
    inline VecXd solveSmallestEig(const SprsD& L, const SprsD& M, int nIter = 50) {
        auto& rt = hip::runtime();
        const suitesparse::detail::DeviceMatrix<double> dL(L, rt.igm_opts), dM(M, rt.igm_opts);
        VecXd u = random_real_vector(L.rows());
        VecXd x = u;
        for (int i = 0; i < nIter; i++) {
            dL.solve(x, dM.apply(u), rt.igm_opts, &rt.last_field);
            const VecXd Mx = dM.apply(x);
            double s = 0.0;
            for (int k = 0; k < x.size(); k++) s += x[k] * Mx[k];
            s = std::sqrt(s);
            for (int k = 0; k < x.size(); k++) u[k] = x[k] / s;
        }
        return x;
    }
}
#endif
