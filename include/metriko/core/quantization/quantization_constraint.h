// metriko (MI355X build) -- QGP quantization: opposite sides of every T-quad have equal
// integer length (reference core/quantization/quantization_constraint.h): row 2 q holds
// +1 on side 0 and -1 on side 2 of T-quad q, row 2 q + 1 the same for sides 1 and 3.
//
// Restated from komietty/metriko include/metriko/core/quantization/quantization_constraint.h
// (Copyright (C) 2025 Saki Komikado), which is subject to the terms of the Mozilla Public License v.2.0.
#ifndef METRIKO_QUANTIZATION_CONSTRAINT_H
#define METRIKO_QUANTIZATION_CONSTRAINT_H
#include <stdexcept>
#include "metriko/core/tmesh/tmesh.h"
namespace metriko {
    inline MatXd compute_constraint(const Tmesh& tmesh) {
        MatXd M = MatXd::Zero(int(tmesh.nTQ) * 2, int(tmesh.nTE));
        for (int iq = 0; iq < int(tmesh.nTQ); iq++) {
            const Tquad& tq = tmesh.tquads[size_t(iq)];
            for (size_t ih = 0; ih < tq.thids.size(); ih++) {
                const int te = tmesh.thalfs[size_t(tq.thids[ih])].edge().id;
                switch (tq.sides[ih]) {
                    case 0: M(iq * 2 + 0, te) = 1; break;
                    case 2: M(iq * 2 + 0, te) = -1; break;
                    case 1: M(iq * 2 + 1, te) = 1; break;
                    case 3: M(iq * 2 + 1, te) = -1; break;
                    default: throw std::invalid_argument("side has to be 0 ~ 3");
                }
            }
        }
        return M;
    }
}
#endif
