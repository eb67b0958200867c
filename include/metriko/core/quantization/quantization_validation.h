// metriko (MI355X build) -- QGP quantization: validity of integer T-edge lengths X
// (reference core/quantization/quantization_validation.h): every length positive, and from
// each half-edge that leaves a singular vertex, the zero-length T-edges reachable through
// adj_thalfs never lead back against the start of a curve.
//
// Restated from komietty/metriko include/metriko/core/quantization/quantization_validation.h
// (Copyright (C) 2025 Saki Komikado), which is subject to the terms of the Mozilla Public License v.2.0.
#ifndef METRIKO_QUANTIZATION_VALIDATION_H
#define METRIKO_QUANTIZATION_VALIDATION_H
#include <queue>
#include "metriko/core/tmesh/tmesh.h"
namespace metriko {
    inline bool compute_validation(const Tmesh& tm, const VecXd& X, const int thid) {
        std::queue<int> q;
        std::vector<bool> visited(tm.nTE, false);
        q.push(tm.thalfs[size_t(thid)].id);
        while (!q.empty()) {
            const int t = q.front();
            q.pop();
            const Thalf& th = tm.thalfs[size_t(t)];
            if (!th.cannonical && th.edge().seg_fr.id == 0) return false;   // the first segment of a curve
            for (const Thalf& pair : th.adj_thalfs()) {
                const int te = pair.edge().id;
                if (!visited[size_t(te)] && X[te] == 0) {
                    visited[size_t(te)] = true;
                    q.push(pair.id);
                }
            }
        }
        return true;
    }
    inline bool compute_validation(const Tmesh& tm, const VecXd& X) {
        for (int i = 0; i < X.size(); i++) if (!(X[i] > 0)) return false;
        for (const Thalf& th : tm.thalfs)
            if (tm.th2sing[th.id] > -1 && !compute_validation(tm, X, th.id)) return false;
        return true;
    }
}
#endif
