// metriko (MI355X build) -- iterative rounding driver.
//
// Drop-in for the reference's core/igm/iterative_rounding/iter_rounding.h:33-113
// (NaiveIntegration KKT, naive LM pass, SeamlessIntegration rounding loop).  The whole
// loop runs behind mk_iterative_rounding: the constraint kernel basis and operator
// assembly on the host, the Poisson KKT, the LM directions, the energies and the FP64
// rounding decisions on the GPU; with localInjectivity the damped barrier LM passes
// (injective_barrier.h, levenberg_marquardt.h DiagonalDamping) run there too.
#ifndef METRIKO_ITER_ROUNDING_H
#define METRIKO_ITER_ROUNDING_H
#include <functional>
#include "metriko/core/common/sparse.h"
#include "metriko/hip/runtime.h"

namespace metriko {
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
    // current fullx after every naive LM iteration and ends that pass when it returns true
    // (iter_rounding_init.h:28); the device loop makes no per-iteration host callbacks, so
    // iter_cb is not called here.
    inline bool iterative_rounding(const VecXi& fixedIdcs, const VecXd& fixedVals, const VecXi& singularIdcs,
                                   const VecXi& integerIdcs, const double length, const SprsD& Cfull, const SprsD& G2,
                                   const int N, const int n, const int nF, const bool seamless, const bool roundSeams,
                                   const bool localInjectivity, const bool verbose, const VecXd& intrinsicField,
                                   const std::function<bool(const VecXd& x)>& iter_cb, VecXd& fullx) {
        (void)iter_cb;
        return iterative_rounding(fixedIdcs, fixedVals, singularIdcs, integerIdcs, length, Cfull, G2, N, n, nF,
                                  seamless, roundSeams, localInjectivity, verbose, intrinsicField, fullx);
    }
}

#endif
