// metriko (MI355X build) -- QEx q-vertices (reference core/qex/gen_q_vert.h:7-56): integer
// uv points on mesh vertices (the first corner's uv within EPS of the grid; its uv is stored
// as the fractional parts), strictly inside mesh edges (collinear, 0 < t < 1 by EPS) and
// strictly inside faces, with their surface positions.
//
// Restated from komietty/metriko include/metriko/core/qex/gen_q_vert.h
// (the komietty/metriko authors), which is subject to the terms of the Mozilla Public License v.2.0.
#ifndef METRIKO_QEX_GEN_Q_VERT_H
#define METRIKO_QEX_GEN_Q_VERT_H
#include "common.h"

namespace metriko::qex {
    inline void generate_q_vert(const Hmesh& mesh, const VecXc& cfn, std::vector<Qvert>& vqvs,
                                std::vector<Qvert>& eqvs, std::vector<Qvert>& fqvs) {
        for (int v = 0; v < mesh.nV; v++) {
            const complex uv = cfn(mesh.crnr[mesh.next[mesh.vert2half[v]]]);
            const double x = std::fmod(std::abs(uv.real()), 1.), y = std::fmod(std::abs(uv.imag()), 1.);
            if ((x < EPS || 1 - x < EPS) && (y < EPS || 1 - y < EPS)) vqvs.emplace_back(complex(x, y), row3(mesh.pos, v), v);
        }
        for (int e = 0; e < mesh.nE; e++) {
            const int h = mesh.edge2half[e];
            const Row3d p1 = row3(mesh.pos, mesh.tail[h]), p2 = row3(mesh.pos, mesh.head[h]);
            const complex uv1 = cfn(mesh.crnr[mesh.next[h]]), uv2 = cfn(mesh.crnr[mesh.prev[h]]);
            const minmax_int b = get_minmax_int({uv1, uv2});
            for (int x = b.min_x; x <= b.max_x; x++)
                for (int y = b.min_y; y <= b.max_y; y++) {
                    const complex xy(x, y);
                    const double a = std::abs(xy - uv1) / std::abs(uv2 - uv1);
                    if (is_collinear(uv1, uv2, xy) && a > EPS && a < 1 - EPS) eqvs.emplace_back(xy, p1 + (p2 - p1) * a, e);
                }
        }
        for (int f = 0; f < mesh.nF; f++) {
            const Row3d p1 = row3(mesh.pos, mesh.idx(f, 0)), p2 = row3(mesh.pos, mesh.idx(f, 1)),
                        p3 = row3(mesh.pos, mesh.idx(f, 2));
            const complex uv1 = cfn(f * 3 + 0), uv2 = cfn(f * 3 + 1), uv3 = cfn(f * 3 + 2);
            const minmax_int b = get_minmax_int({uv1, uv2, uv3});
            for (int x = b.min_x; x <= b.max_x; x++)
                for (int y = b.min_y; y <= b.max_y; y++) {
                    const complex xy(x, y);
                    if (is_inside_triangle(uv1, uv2, uv3, xy))
                        fqvs.emplace_back(xy, conversion_2d_3d(uv1, uv2, uv3, p1, p2, p3, xy), f);
                }
        }
    }
}
#endif
