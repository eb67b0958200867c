// metriko (MI355X build) -- QGP quantization: generating vectors
// (reference core/quantization/quantization_evaluation.h).  One basis loop per canonical
// T-half-edge; row i counts the T-edges of loop i; a maximal linearly independent subset of
// the rows (reduce_to_linearly_independent) transposed gives the generators as columns.
//
// Restated from komietty/metriko include/metriko/core/quantization/quantization_evaluation.h
// (Copyright (C) 2025 Saki Komikado), which is subject to the terms of the Mozilla Public License v.2.0.
#ifndef METRIKO_QUANTIZATION_EVALUATION_H
#define METRIKO_QUANTIZATION_EVALUATION_H
#include "quantization_basisloop.h"
#include "metriko/core/common/sparse.h"
#include "metriko/core/solver/matrix_ops.h"
namespace metriko {
    template <typename Func>
    MatXd construct_generating_vectors(const Tmesh& tmesh, const VecXd& R, Func compare) {
        std::vector<TripD> T;
        int irow = 0;
        for (const Thalf& th : tmesh.thalfs) {
            if (!th.cannonical) continue;
            for (const int t : gen_basis_loop(tmesh.tquads, tmesh.thalfs, tmesh.th2quad, tmesh.th2side, R, th, compare))
                T.emplace_back(irow, tmesh.thalfs[size_t(t)].teid, 1.0);
            irow++;
        }
        SprsD G(int(tmesh.nTE), int(tmesh.nTE));
        G.setFromTriplets(T);
        reduce_to_linearly_independent(G);
        MatXd D = MatXd::Zero(G.cols(), G.rows());   // generators as columns
        for (int i = 0; i < G.rows(); i++)
            for (int64_t q = G.ptr[i]; q < G.ptr[i + 1]; q++) D(G.col[q], i) = G.val[q];
        return D;
    }
}
#endif
