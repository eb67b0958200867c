// metriko (MI355X build) -- iterative rounding driver.
//
// Drop-in for the reference's core/igm/iterative_rounding/iter_rounding.h:33-113
// (NaiveIntegration KKT, naive LM pass, SeamlessIntegration rounding loop).  The whole
// loop runs behind mk_iterative_rounding: the constraint kernel basis and operator
// assembly on the host, the Poisson KKT, the LM directions, the energies and the FP64
// rounding decisions on the GPU; with localInjectivity the damped barrier LM passes
// (injective_barrier.h, levenberg_marquardt.h DiagonalDamping) run there too.
//
// Restated from komietty/metriko include/metriko/core/igm/iterative_rounding/iter_rounding.h
// (Copyright (C) 2025 Saki Komikado), which is subject to the terms of the Mozilla Public License v.2.0.
#ifndef METRIKO_ITER_ROUNDING_H
#define METRIKO_ITER_ROUNDING_H
#include <functional>
#include <memory>
#include "metriko/core/common/sparse.h"
#include "metriko/core/solver/suitesparse/suitesparse_pdefinite.h"
#include "metriko/hip/runtime.h"

namespace metriko {
    // CholeskyWrapper (iter_rounding.h:14-31): the LinearSolver of LMSolver.  factorize()
    // uploads the SPD normal matrix to the device, solve() runs multigrid PCG there (FP64 relative
    // residual 1e-10 by default); failures are reported as false, which LMSolver turns into
    // "Solver Failed to factorize!" (levenberg_marquardt.h:184-187).
    class CholeskyWrapper {
    public:
        std::unique_ptr<suitesparse::PositiveDefiniteSolver<double>> llt;

        bool factorize(const SprsD& A) {
            try {
                llt = std::make_unique<suitesparse::PositiveDefiniteSolver<double>>(A);
                return true;
            } catch (const std::exception&) {
                llt.reset();
                return false;
            }
        }
        bool solve(const VecXd& rhs, VecXd& x) const {
            if (!llt) return false;
            try {
                llt->solve(x, rhs);
                return true;
            } catch (const std::exception&) {
                return false;
            }
        }
    };

    inline bool iterative_rounding(const VecXi& fixedIdcs, const VecXd& fixedVals, const VecXi& singularIdcs,
                                   const VecXi& integerIdcs, const double length, const SprsD& Cfull, const SprsD& G2,
                                   const int N, const int n, const int nF, const bool seamless, const bool roundSeams,
                                   const bool localInjectivity, const bool verbose, const VecXd& intrinsicField,
                                   VecXd& fullx, std::vector<int>* decIndex = nullptr,
                                   std::vector<int>* decValue = nullptr) {
        auto& rt = hip::runtime();
        mk_igm_flags fl{seamless ? 1 : 0, roundSeams ? 1 : 0, localInjectivity ? 1 : 0, verbose ? 1 : 0};
        fullx = VecXd(G2.cols());
        const int cap = singularIdcs.size() + integerIdcs.size() + 1;
        std::vector<int32_t> di(static_cast<size_t>(cap)), dv(static_cast<size_t>(cap));
        std::vector<double> dm(static_cast<size_t>(cap));
        int ndec = 0, success = 0;
        hip::check(mk_iterative_rounding(rt.get(), fixedIdcs.size(), fixedIdcs.data(), fixedVals.data(),
                                         singularIdcs.size(), singularIdcs.data(), integerIdcs.size(), integerIdcs.data(),
                                         length, Cfull.rows(), Cfull.cols(), Cfull.ptr.data(), Cfull.col.data(),
                                         Cfull.val.data(), G2.rows(), G2.cols(), G2.ptr.data(), G2.col.data(),
                                         G2.val.data(), N, n, nF, &fl, intrinsicField.data(), &rt.igm_opts,
                                         fullx.data(), &success, cap, di.data(), dv.data(), dm.data(), &ndec,
                                         &rt.last_igm),
                   "mk_iterative_rounding");
        if (decIndex) decIndex->assign(di.begin(), di.begin() + ndec);
        if (decValue) decValue->assign(dv.begin(), dv.begin() + ndec);
        return success != 0;
    }

    // Reference signature (iter_rounding.h:33-51).  The reference calls iter_cb with the
    // current fullx (UFull x) after every naive LM iteration and ends that pass when it returns
    // true (iter_rounding_init.h:28, levenberg_marquardt.h:221-224): a non-empty iter_cb is
    // installed as the context's naive-pass callback for this call and cleared afterwards.
    inline bool iterative_rounding(const VecXi& fixedIdcs, const VecXd& fixedVals, const VecXi& singularIdcs,
                                   const VecXi& integerIdcs, const double length, const SprsD& Cfull, const SprsD& G2,
                                   const int N, const int n, const int nF, const bool seamless, const bool roundSeams,
                                   const bool localInjectivity, const bool verbose, const VecXd& intrinsicField,
                                   const std::function<bool(const VecXd& x)>& iter_cb, VecXd& fullx) {
        if (!iter_cb)
            return iterative_rounding(fixedIdcs, fixedVals, singularIdcs, integerIdcs, length, Cfull, G2, N, n, nF,
                                      seamless, roundSeams, localInjectivity, verbose, intrinsicField, fullx);
        auto& rt = hip::runtime();
        struct Call {
            const std::function<bool(const VecXd&)>* cb;
            static int trampoline(void* user, const double* x, int nx) {
                const auto* self = static_cast<const Call*>(user);
                VecXd v(nx);
                for (int i = 0; i < nx; i++) v[i] = x[i];
                return (*self->cb)(v) ? 1 : 0;
            }
        } call{&iter_cb};
        struct Clear {
            mk_ctx* c;
            ~Clear() { mk_set_naive_iter_callback(c, nullptr, nullptr); }
        } clear{rt.get()};
        hip::check(mk_set_naive_iter_callback(rt.get(), &Call::trampoline, &call), "mk_set_naive_iter_callback");
        return iterative_rounding(fixedIdcs, fixedVals, singularIdcs, integerIdcs, length, Cfull, G2, N, n, nF,
                                  seamless, roundSeams, localInjectivity, verbose, intrinsicField, fullx);
    }
}

#endif
