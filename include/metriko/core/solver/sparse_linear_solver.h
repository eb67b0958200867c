// metriko (MI355X build) -- solver seam of the field stage.
//
// Drop-in for the reference's core/solver/sparse_linear_solver.h.  Where the
// reference factorises L once (CHOLMOD simplicial LDL^H or Eigen SparseLU,
// sparse_linear_solver.h:29-54) and back-substitutes nIter times, this build
// assembles L on the GPU and runs nIter warm-started PCG solves there
// (FP32 inner + FP64 refinement by default), with the normalisation and
// Rayleigh steps in FP64.
#ifndef METRIKO_SPARSE_LINEAR_SOLVER_H
#define METRIKO_SPARSE_LINEAR_SOLVER_H
#include <cstdlib>
#include "metriko/core/common/sparse.h"
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

    // solveSquare (sparse_linear_solver.h:16-27) for the Hermitian positive definite lhs the
    // reference's callers pass (CurvatureAligned: L - 0*M): PCG on the device.
    inline VecXc solveSquare(SprsC& lhs, VecXc& rhs) {
        auto& rt = hip::runtime();
        DeviceOp d = upload_complex(lhs);
        VecXc x = VecXc::Zero(lhs.rows());
        hip::check(mk_spd_solve(rt.get(), d.op, reinterpret_cast<const double*>(rhs.data()),
                                reinterpret_cast<double*>(x.data()), &rt.field_opts, &rt.last_field),
                   "mk_spd_solve");
        return x;
    }
}
#endif
