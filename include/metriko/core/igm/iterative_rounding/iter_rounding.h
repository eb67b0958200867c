// metriko (MI355X build) -- iterative rounding driver.
//
// Drop-in for the reference's core/igm/iterative_rounding/iter_rounding.h:33-113
// (NaiveIntegration KKT, naive LM pass, SeamlessIntegration rounding loop).  The whole
// loop runs behind mk_iterative_rounding: the constraint kernel basis and operator
// assembly on the host, the Poisson KKT, the LM directions, the energies and the FP64
// rounding decisions on the GPU.  Local injectivity (the barrier LM) is not built in
// this tier: requesting it raises.
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
        if (localInjectivity) throw std::runtime_error("iterative_rounding: localInjectivity is not built in this tier");
        auto& rt = hip::runtime();
        mk_igm_flags fl{seamless ? 1 : 0, roundSeams ? 1 : 0, 0, verbose ? 1 : 0};
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

    // Reference signature (iter_rounding.h:33-51); the callback only acts with local injectivity.
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
