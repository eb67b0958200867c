// metriko (MI355X build) -- the reference's SuiteSparse solver interface, on the GPU.
//
// Drop-in for the reference's core/solver/suitesparse/suitesparse_utilities.h:16-24: the
// LinearSolver<T> base class (the constructor factorises, solve(x, rhs) back-substitutes).
// Here the constructor uploads the matrix to the device (hybrid SELL operator, FP64 and FP32
// value copies) and solve() runs PCG there (mk_spd_solve): FP64 Krylov loop with an FP32
// multigrid V-cycle by default, to a 1e-10 FP64 relative residual.  Failures throw
// std::runtime_error, as the reference's CHOLMOD / UMFPACK wrappers do.
//
// Restated from komietty/metriko include/metriko/core/solver/suitesparse/suitesparse_utilities.h
// (the komietty/metriko authors), which is subject to the terms of the Mozilla Public License v.2.0.
#ifndef METRIKO_SUITESPARSE_UTILITIES_H
#define METRIKO_SUITESPARSE_UTILITIES_H
#include <complex>
#include <stdexcept>
#include <string>
#include <type_traits>
#include "metriko/core/common/sparse.h"
#include "metriko/hip/runtime.h"

namespace metriko { namespace suitesparse {
    template <typename K> using VecX = metriko::VecX<K>;
    template <typename K> using Sprs = metriko::Csr<K>;

    enum class SType { UNSYMMETRIC = 0, SYMMETRIC };

    template <typename T>
    class LinearSolver {
    public:
        explicit LinearSolver(const Sprs<T>& mat) : nRows(size_t(mat.rows())), nCols(size_t(mat.cols())) {}
        virtual ~LinearSolver() {}
        virtual void solve(VecX<T>& x, const VecX<T>& rhs) = 0;

    protected:
        size_t nRows, nCols;
    };

    namespace detail {
        // Device operator of a square CSR matrix (real or complex), owned.
        template <typename T>
        struct DeviceMatrix {
            mk_op* op = nullptr;
            int n = 0;
            DeviceMatrix() = default;
            DeviceMatrix(const Sprs<T>& A, const mk_solve_opts& o) {
                if (A.rows() != A.cols()) throw std::runtime_error("metriko: solver matrix must be square");
                n = A.rows();
                auto& rt = hip::runtime();
                mk_status s;
                if constexpr (std::is_same<T, double>::value)
                    s = mk_op_create_real(rt.get(), n, A.ptr.data(), A.col.data(), A.val.data(), &o, &op);
                else
                    s = mk_op_create_complex(rt.get(), n, A.ptr.data(), A.col.data(),
                                             reinterpret_cast<const double*>(A.val.data()), &o, &op);
                hip::check(s, "mk_op_create");
            }
            DeviceMatrix(const DeviceMatrix&) = delete;
            DeviceMatrix& operator=(const DeviceMatrix&) = delete;
            DeviceMatrix(DeviceMatrix&& o) noexcept : op(o.op), n(o.n) { o.op = nullptr; }
            DeviceMatrix& operator=(DeviceMatrix&& o) noexcept {
                if (op) mk_op_destroy(op);
                op = o.op; n = o.n; o.op = nullptr;
                return *this;
            }
            ~DeviceMatrix() { if (op) mk_op_destroy(op); }

            // x = A^-1 b by PCG (x: warm start in, solution out)
            void solve(VecX<T>& x, const VecX<T>& b, const mk_solve_opts& o, mk_stats* st) const {
                if (b.size() != n) throw std::runtime_error("metriko: right-hand side size");
                if (x.size() != n) x = VecX<T>::Zero(n);
                hip::check(mk_spd_solve(hip::runtime().get(), op, reinterpret_cast<const double*>(b.data()),
                                        reinterpret_cast<double*>(x.data()), &o, st),
                           "mk_spd_solve");
            }
            // y = A x
            VecX<T> apply(const VecX<T>& x) const {
                VecX<T> y(n);
                hip::check(mk_op_apply(hip::runtime().get(), op, reinterpret_cast<const double*>(x.data()),
                                       reinterpret_cast<double*>(y.data())),
                           "mk_op_apply");
                return y;
            }
        };

        template <typename T> inline T conj_of(const T& v) { return v; }
        template <> inline std::complex<double> conj_of(const std::complex<double>& v) { return std::conj(v); }
        inline double abs_of(double v) { return std::abs(v); }
        inline double abs_of(const std::complex<double>& v) { return std::abs(v); }

        // A^H (conjugate transpose), rows sorted
        template <typename T>
        Sprs<T> adjoint(const Sprs<T>& A) {
            Sprs<T> H = A.transpose();
            for (auto& v : H.val) v = conj_of(v);
            return H;
        }

        // A == A^H to rel_tol of the largest entry (same pattern required)
        template <typename T>
        bool is_hermitian(const Sprs<T>& A, double rel_tol = 1e-14) {
            if (A.rows() != A.cols()) return false;
            const Sprs<T> H = adjoint(A);
            if (H.ptr != A.ptr || H.col != A.col) return false;
            double amax = 0.0, dmax = 0.0;
            for (size_t k = 0; k < A.val.size(); k++) {
                amax = std::max(amax, abs_of(A.val[k]));
                dmax = std::max(dmax, abs_of(A.val[k] - H.val[k]));
            }
            return dmax <= rel_tol * amax;
        }
    }  // namespace detail
}}  // namespace metriko::suitesparse

#endif
