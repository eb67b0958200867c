// metriko (MI355X build) -- QEx sanitization (reference core/qex/sanitization.h:9-57).
//
// Per vertex, in vertex order: the first corner around it (Vert::adjHalfs order) is snapped
// -- to the integer grid at a singular vertex, else through x + d - d with d the power of
// two of the largest |u|, |v| around the vertex, which rounds every corner of the vertex to
// the same binary grid -- and the other corners are rebuilt from it through the transitions
// of the halfedges crossed, so corners of one vertex agree exactly across seams.
//
// Restated from komietty/metriko include/metriko/core/qex/sanitization.h
// (the komietty/metriko authors), which is subject to the terms of the Mozilla Public License v.2.0.
#ifndef METRIKO_QEX_SANITIZATION_H
#define METRIKO_QEX_SANITIZATION_H
#include "common.h"

namespace metriko::qex {
    inline void sanitization(const Hmesh& mesh, const VecXi& matching, const VecXi& singular, const int rosyN,
                             VecXc& cfn) {
        require_closed(mesh);
        VecXc heR, heT;
        compute_trs_matrix(mesh, cfn, matching, rosyN, heR, heT);
        for (int v = 0; v < mesh.nV; v++) {
            const std::vector<int> hs = vert_halfs(mesh, v);
            double mx = 0;
            for (int h : hs) {
                const complex uv = cfn(mesh.crnr[mesh.next[h]]);
                mx = std::max(mx, std::abs(uv.real()));
                mx = std::max(mx, std::abs(uv.imag()));
            }
            for (size_t i = 0; i < hs.size(); i++) {
                const int h = hs[i];
                const int cc = mesh.crnr[mesh.next[h]];
                if (i == 0) {
                    const complex uv = cfn(cc);
                    if (singular[v]) {
                        cfn(cc) = complex(std::round(uv.real()), std::round(uv.imag()));
                    } else {
                        const double delta = std::pow(2, std::log2(mx));
                        const complex sign((0 < uv.real()) - (uv.real() < 0), (0 < uv.imag()) - (uv.imag() < 0));
                        cfn(cc) = (cfn(cc) + delta * sign) - delta * sign;
                    }
                } else {
                    const int t = mesh.twin[h];
                    cfn(cc) = heR(t) * cfn(mesh.crnr[mesh.prev[t]]) + heT(t);
                }
            }
        }
    }
}
#endif
