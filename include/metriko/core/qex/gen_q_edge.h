// metriko (MI355X build) -- QEx edges (reference core/qex/gen_q_edge.h:7-119).
//
// Every unconnected port, in port order, traces its grid line through the faces: in a face
// whose interior holds the next grid point it pairs with the first face port there facing
// back; else with the first unconnected edge port (then vertex port) on the face's edges
// (vertices) whose 3D direction is collinear with the trace and opposite to it (EPS on the
// cosines); else it crosses the exit edge (pick_next_half) into the twin face through the
// transition (heR, heT).  The reference scans the whole port list for each test; here the
// candidates come from per-face / per-edge / per-vertex lists in port order, which gives the
// same first match.  A trace longer than 4 nF + 16 faces throws (the reference would loop).
//
// Restated from komietty/metriko include/metriko/core/qex/gen_q_edge.h
// (the komietty/metriko authors), which is subject to the terms of the Mozilla Public License v.2.0.
#ifndef METRIKO_QEX_GEN_Q_EDGE_H
#define METRIKO_QEX_GEN_Q_EDGE_H
#include <string>
#include "common.h"

namespace metriko::qex {
    inline bool predict_extrinsic_collinear(const Hmesh& mesh, const VecXc& cfn, complex ori, complex dir, Face face,
                                            const Qport& pair) {
        const Face pf{{pair.fid, &mesh}};
        const Row3d pb1 = conversion_2d_3d(face, cfn, ori);
        const Row3d pb2 = conversion_2d_3d(face, cfn, ori + dir);
        const Row3d pa1 = conversion_2d_3d(pf, cfn, pair.uv);
        const Row3d pa2 = conversion_2d_3d(pf, cfn, pair.uv + pair.dir);
        const Row3d da = (pa2 - pa1).normalized();
        const Row3d db = (pb2 - pb1).normalized();
        const Row3d dc = (pa1 - pb1).normalized();
        return std::abs(1. - db.dot(dc)) < EPS && std::abs(1. + da.dot(db)) < EPS;
    }

    // the halfedge of `face` the ray ori + s dir leaves through, and the exit uv
    inline std::pair<int, complex> pick_next_half(const Hmesh& mesh, const VecXc& cfn, complex ori, complex dir, int face) {
        int hs[3];
        face_halfs(mesh, face, hs);
        for (const int h : hs) {
            const complex uv1 = cfn(mesh.crnr[mesh.next[h]]), uv2 = cfn(mesh.crnr[mesh.prev[h]]);
            double rab, rcd;
            if (find_strict_intersection(ori, ori + dir * 1e2, uv1, uv2, rab, rcd) && rab > EPS)
                return {h, lerp(uv1, uv2, rcd)};
        }
        throw std::runtime_error("qex: no next half found");
    }

    inline std::vector<Qedge> generate_q_edge(const Hmesh& mesh, const VecXc& cfn, const VecXi& matching,
                                              std::vector<Qport>& qports) {
        require_closed(mesh);
        std::vector<Qedge> qedges;
        VecXc heR, heT;
        compute_trs_matrix(mesh, cfn, matching, 4, heR, heT);
        std::vector<std::vector<int>> byFace(size_t(mesh.nF)), byEdge(size_t(mesh.nE)), byVert(size_t(mesh.nV));
        for (const Qport& p : qports) {
            if (p.fid >= 0) byFace[size_t(p.fid)].push_back(p.idx);
            if (p.eid >= 0) byEdge[size_t(p.eid)].push_back(p.idx);
            if (p.vid >= 0) byVert[size_t(p.vid)].push_back(p.idx);
        }
        const int maxSteps = 4 * mesh.nF + 16;
        for (Qport& pfr : qports) {
            if (pfr.isConnected) continue;
            complex ori = pfr.uv, dir = pfr.dir;
            complex gri = nearby_grid(ori, dir);
            int f = pfr.fid;
            Qport* hit = nullptr;
            for (int step = 0; step < maxSteps && !hit; step++) {
                if (is_inside_triangle(cfn(f * 3), cfn(f * 3 + 1), cfn(f * 3 + 2), gri)) {
                    for (const int q : byFace[size_t(f)]) {
                        const Qport& p = qports[size_t(q)];
                        if (p.idx != pfr.idx && equal(p.dir, -dir) && std::abs(p.uv - gri) < EPS) { hit = &qports[size_t(q)]; break; }
                    }
                    if (!hit) throw std::runtime_error("qex: port " + std::to_string(pfr.idx) + " finds no face port");
                    break;
                }
                int hs[3];
                face_halfs(mesh, f, hs);
                const Face face{{f, &mesh}};
                for (const int h : hs) {
                    for (const int q : byEdge[size_t(mesh.edge[h])]) {
                        const Qport& p = qports[size_t(q)];
                        if (p.isConnected || p.eid == pfr.eid) continue;
                        if (predict_extrinsic_collinear(mesh, cfn, ori, dir, face, p)) { hit = &qports[size_t(q)]; break; }
                    }
                    if (hit) break;
                }
                if (hit) break;
                for (const int h : hs) {
                    for (const int q : byVert[size_t(mesh.tail[h])]) {
                        const Qport& p = qports[size_t(q)];
                        if (p.isConnected || p.vid == pfr.vid) continue;
                        if (predict_extrinsic_collinear(mesh, cfn, ori, dir, face, p)) { hit = &qports[size_t(q)]; break; }
                    }
                    if (hit) break;
                }
                if (hit) break;
                const auto [nh, x] = pick_next_half(mesh, cfn, ori, dir, f);
                f = mesh.face[mesh.twin[nh]];
                ori = heR(nh) * x + heT(nh);
                dir = heR(nh) * dir;
                gri = nearby_grid(ori, dir);
            }
            if (!hit) throw std::runtime_error("qex: port " + std::to_string(pfr.idx) + " traced too far without a pair");
            pfr.isConnected = true;
            hit->isConnected = true;
            qedges.emplace_back(pfr, *hit);
        }
        return qedges;
    }
}
#endif
