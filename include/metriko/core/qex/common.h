// metriko (MI355X build) -- QEx quad extraction: types and predicates.
//
// Drop-in for the reference's core/qex/common.h: Qvert / Qport / Qedge / Qhalf / Qface,
// compute_trs_matrix (the per-halfedge rotation r = i^m and integer translation t of the
// transition from the face of h to the face of its twin, qex/common.h:73-94) and the
// orientation / grid predicates with EPS = 1e-9.  The extraction runs on the host, on the
// corner uv the GPU parameterization produced (cfn, corner 3 f + j at vertex F(f, j)).
//
// Restated from komietty/metriko include/metriko/core/qex/common.h
// (the komietty/metriko authors), which is subject to the terms of the Mozilla Public License v.2.0.
#ifndef METRIKO_QEX_COMMON_H
#define METRIKO_QEX_COMMON_H
#include <cmath>
#include <stdexcept>
#include <vector>
#include "metriko/core/common/utilities.h"
#include "metriko/core/hmesh/hmesh.h"
#include "metriko/core/hmesh/utilities.h"

namespace metriko::qex {
    struct Qvert {
        complex uv;
        Row3d pos;
        int sid;   // vertex, edge or face id the q-vertex lies on
        Qvert(const complex& uv_, const Row3d& pos_, const int sid_) : uv(uv_), pos(pos_), sid(sid_) {}
    };

    struct Qport {
        int idx, vid, eid, fid;
        complex uv, dir;
        Row3d pos;
        bool isConnected = false;
        int next_id = -1;
        int prev_id = -1;
        Qport(int idx_, int vid_, int eid_, int fid_, complex uv_, complex dir_, const Row3d& pos_)
            : idx(idx_), vid(vid_), eid(eid_), fid(fid_), uv(uv_), dir(dir_), pos(pos_) {}
    };

    struct Qedge {
        const Qport& port1;
        const Qport& port2;
        Qedge(const Qport& p1, const Qport& p2) : port1(p1), port2(p2) {}
    };

    struct Qhalf {
        const Qedge& edge;
        int idx;
        bool cannonical;
        const Qport& port1() const { return cannonical ? edge.port1 : edge.port2; }
        const Qport& port2() const { return cannonical ? edge.port2 : edge.port1; }
        Qhalf(const Qedge& e, int i, bool c) : edge(e), idx(i), cannonical(c) {}
    };

    struct Qface { std::vector<Qhalf> qhalfs; };

    inline void compute_trs_matrix(const Hmesh& mesh, const VecXc& cfn, const VecXi& matching, const int rosyN,
                                   VecXc& heR, VecXc& heT) {
        heR.resize(mesh.nH);
        heT.resize(mesh.nH);
        for (int h = 0; h < mesh.nH; h++) {
            const complex uv1 = cfn(mesh.crnr[mesh.next[h]]);
            const complex uv2 = cfn(mesh.crnr[mesh.prev[mesh.twin[h]]]);
            int m = (mesh.edge2half[mesh.edge[h]] == h ? 1 : -1) * matching[mesh.edge[h]];
            m = m < 0 ? (rosyN + m % rosyN) % rosyN : m % rosyN;
            const complex r = get_quater_rot(m);
            const complex t = uv2 - r * uv1;
            heR(h) = r;
            heT(h) = complex(std::round(t.real()), std::round(t.imag()));
        }
    }

    inline double orientation(complex pa, complex pb, complex pc) {
        const double acx = pa.real() - pc.real(), bcx = pb.real() - pc.real();
        const double acy = pa.imag() - pc.imag(), bcy = pb.imag() - pc.imag();
        return acx * bcy - acy * bcx;
    }
    inline bool is_collinear(complex pa, complex pb, complex pc) { return std::abs(orientation(pa, pb, pc)) < EPS; }
    inline bool is_points_into(complex p1, complex p2, complex p3, complex uv) {
        return orientation(p1, p2, uv) > EPS && orientation(p1, p3, uv) < EPS;
    }
    inline bool is_inside_triangle(complex pa, complex pb, complex pc, complex uv) {
        return orientation(pa, pb, uv) > EPS && orientation(pb, pc, uv) > EPS && orientation(pc, pa, uv) > EPS;
    }
    inline bool is_inside_triangle(const Face f, const VecXc& cfn, const complex uv) {
        return is_inside_triangle(cfn(f.id * 3 + 0), cfn(f.id * 3 + 1), cfn(f.id * 3 + 2), uv);
    }
    inline complex nearby_grid(const complex a) { return {std::round(a.real()), std::round(a.imag())}; }
    inline double nearby_grid(double x, double dir) {
        if (dir == 0) return x;
        const double f = std::abs(std::fmod(x, 1.));
        if (f < EPS || 1 - f < EPS) x += dir * EPS;
        return dir > 0 ? std::ceil(x) : std::floor(x);
    }
    inline complex nearby_grid(complex uv, complex dir) {
        return {nearby_grid(uv.real(), dir.real()), nearby_grid(uv.imag(), dir.imag())};
    }

    // halfedges of a face in Face::adjHalfs order (face2half, next, next)
    inline void face_halfs(const Hmesh& mesh, int f, int out[3]) {
        out[0] = mesh.face2half[f];
        out[1] = mesh.next[out[0]];
        out[2] = mesh.next[out[1]];
    }
    // outgoing halfedges of v in Vert::adjHalfs order (from vert2half, h -> h.prev().twin())
    inline std::vector<int> vert_halfs(const Hmesh& mesh, int v) {
        std::vector<int> hs;
        mesh.forVertCcw(mesh.vert2half[v], [&](int h) { hs.push_back(h); return true; });
        return hs;
    }
    inline void require_closed(const Hmesh& mesh) {
        for (int h = 0; h < mesh.nH; h++)
            if (mesh.face[h] < 0) throw std::runtime_error("qex: closed meshes only");
    }
}
#endif
