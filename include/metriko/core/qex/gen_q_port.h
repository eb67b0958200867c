// metriko (MI355X build) -- QEx ports (reference core/qex/gen_q_port.h:6-107): the
// outgoing grid directions of every q-vertex, linked cyclically (prev / next) per q-vertex.
// Vertex q-verts: per face around the vertex, the directions pointing into that face's
// corner (or along its first edge); edge q-verts: the four directions across the two faces
// of the edge, deduplicated by the normalised surface point one unit away; face q-verts:
// the four axis directions.
//
// Restated from komietty/metriko include/metriko/core/qex/gen_q_port.h
// (the komietty/metriko authors), which is subject to the terms of the Mozilla Public License v.2.0.
#ifndef METRIKO_QEX_GEN_Q_PORT_H
#define METRIKO_QEX_GEN_Q_PORT_H
#include "common.h"

namespace metriko::qex {
    namespace detail {
        inline void link_ports(std::vector<Qport>& tmp, std::vector<Qport>& qports) {
            const int s = int(tmp.size());
            for (int i = 0; i < s; i++) tmp[i].idx = int(qports.size()) + i;
            for (int i = 0; i < s; i++) {
                tmp[i].prev_id = tmp[(i - 1 + s) % s].idx;
                tmp[i].next_id = tmp[(i + 1 + s) % s].idx;
            }
            qports.insert(qports.end(), tmp.begin(), tmp.end());
        }
        // first r with uv1 + i^r pointing into the corner (uv1; uv2, uv3); 4 if none
        inline int first_inward(complex uv1, complex uv2, complex uv3) {
            int r = 0;
            while (r < 4 && !is_points_into(uv1, uv2, uv3, uv1 + get_quater_rot(r))) r++;
            return r;
        }
    }

    inline void generate_eqvert_qport(const Hmesh& mesh, const VecXc& cfn, const std::vector<Qvert>& eqverts,
                                      std::vector<Qport>& qports) {
        std::vector<Qport> tmp;
        std::vector<Row3d> visit;
        for (const Qvert& qv : eqverts) {
            tmp.clear();
            const int h0 = mesh.edge2half[qv.sid];
            for (const int h : {h0, mesh.twin[h0]}) {
                const int c1 = mesh.crnr[mesh.next[h]], c2 = mesh.crnr[mesh.prev[h]];
                const complex uv1 = cfn(c1), uv2 = cfn(c2), uv3 = cfn(mesh.crnr[h]);
                const Row3d p1 = row3(mesh.pos, mesh.tail[h]), p2 = row3(mesh.pos, mesh.head[h]);
                const complex uv = lerp(uv1, uv2, (qv.pos - p1).norm() / (p2 - p1).norm());
                const int r = detail::first_inward(uv1, uv2, uv3);
                const Face face{{mesh.face[h], &mesh}};
                for (int i = 0; i < 4; i++) {
                    const complex dir = get_quater_rot((r + i + 2) % 4);
                    const bool f1 = orientation(uv1, uv2, uv + dir) >= 0;
                    const Row3d ev = conversion_2d_3d(face, cfn, uv + dir).normalized();
                    bool seen = false;
                    for (const Row3d& w : visit) if ((ev - w).norm() < EPS) { seen = true; break; }
                    if (f1 && !seen) {
                        tmp.emplace_back(-1, -1, qv.sid, mesh.face[h], nearby_grid(uv), dir, qv.pos);
                        visit.push_back(ev);
                    }
                }
            }
            // the reference asserts 4 ports here (gen_q_port.h:47), compiled out in release
            // builds: the ports found are linked as they are (`visit` is shared by all edge
            // q-vertices, so a probe point another q-vertex already used is skipped)
            detail::link_ports(tmp, qports);
        }
    }

    inline void generate_fqvert_qport(const Hmesh& mesh, const std::vector<Qvert>& fqverts, std::vector<Qport>& qports) {
        (void)mesh;
        for (const Qvert& qv : fqverts) {
            for (int i = 0; i < 4; i++)
                qports.emplace_back(int(qports.size()), -1, -1, qv.sid, qv.uv, get_quater_rot(i), qv.pos);
            const int l = int(qports.size());
            for (int i = 0; i < 4; i++) {
                qports[l - i - 1].next_id = qports[l - (i - 1 + 4) % 4 - 1].idx;
                qports[l - i - 1].prev_id = qports[l - (i + 1 + 4) % 4 - 1].idx;
            }
        }
    }

    inline void generate_vqvert_qport(const Hmesh& mesh, const VecXc& cfn, const std::vector<Qvert>& vqverts,
                                      std::vector<Qport>& qports) {
        std::vector<Qport> tmp;
        for (const Qvert& qv : vqverts) {
            tmp.clear();
            for (const int h : vert_halfs(mesh, qv.sid)) {
                const complex uv1 = cfn(mesh.crnr[mesh.next[h]]), uv2 = cfn(mesh.crnr[mesh.prev[h]]),
                              uv3 = cfn(mesh.crnr[h]);
                const int r = detail::first_inward(uv1, uv2, uv3);
                for (int i = 0; i < 4; i++) {
                    const complex d = get_quater_rot((r + i + 3) % 4);
                    const bool f1 = is_points_into(uv1, uv2, uv3, uv1 + d);
                    const bool f2 = is_collinear(uv1, uv2, uv1 + d);
                    const bool f3 = dot(uv2 - uv1, d) > 0;
                    if (f1 || (f2 && f3)) tmp.emplace_back(-1, qv.sid, -1, mesh.face[h], uv1, d, qv.pos);
                }
            }
            detail::link_ports(tmp, qports);
        }
    }
}
#endif
