// metriko (MI355X build) -- QGP motorcycle graph on a (non-seamless, locally injective) uv.
//
// Drop-in for the reference's core/tmesh/motorcycle.h (MotorcycleGraph, Mcurv, Msgmt,
// Mvert, Mport), host code downstream of the GPU parameterization:
//  * ports (motorcycle.h:165-212): per singular vertex in ascending id, per outgoing halfedge
//    in Vert::adjHalfs order, the grid directions that point into the corner or run along one
//    of its edges, taken from the first rotation after the first non-inward one, sorted by the
//    dot product with the corner's first edge (descending, stable), numbered in that order and
//    linked cyclically per vertex;
//  * curves (motorcycle.h:214-296): each port's curve grows one face per sweep (all curves in
//    port order) until it strictly crosses a segment of another curve in the face (the
//    nearest crossing); that segment is split there (HitL / HitR by the side the crash comes
//    from) and the curve ends in HitB.  Crossing an edge turns the direction by
//    polar(1, pi/2 m), m the edge's matching seen from the crossing halfedge.
// A curve that has not crashed after 4 nF + 16 sweeps throws (the reference would loop).
//
// Restated from komietty/metriko include/metriko/core/tmesh/motorcycle.h
// (Copyright (C) 2025 Saki Komikado), which is subject to the terms of the Mozilla Public License v.2.0.
#ifndef METRIKO_MOTORCYCLE_H
#define METRIKO_MOTORCYCLE_H
#include <algorithm>
#include <cmath>
#include <stdexcept>
#include <vector>
#include "../common/utilities.h"
#include "../hmesh/hmesh.h"

namespace metriko {
    class Mcurv;
    class MotorcycleGraph;

    enum MvertType { None, HitL, HitR, HitB };

    struct Mport {
        int id;
        Vert vert;
        Face face;
        complex uv, dir;
        int next = -1;
        int prev = -1;
        Mport(int id_, Vert v, Face f, complex uv_, complex dir_) : id(id_), vert(v), face(f), uv(uv_), dir(dir_) {}
    };

    struct Mvert {
        complex uv;
        Mcurv* crash = nullptr;
        MvertType type = None;
    };

    struct Msgmt {
        Mcurv* curv;
        Face face;
        Mvert fr, to;
        int id = -1, prev_id = -1, next_id = -1;
        bool operator==(const Msgmt& o) const { return face.id == o.face.id && fr.uv == o.fr.uv && to.uv == o.to.uv; }
        complex diff() const { return to.uv - fr.uv; }
        const Msgmt& next() const;
        const Msgmt& prev() const;
    };

    struct Cache {
        int half = -1;   // halfedge the curve left its last face through
        double ratio = 0;
        complex dir;
        bool intersected = false;
    };

    class Mcurv {
    public:
        const Mport& port;
        std::vector<Msgmt> sgmts;
        Cache cache;
        explicit Mcurv(const Mport& p) : port(p) {}
        bool operator==(const Mcurv& o) const { return port.id == o.port.id; }

        // split the segment equal to `target` at `ratio` where `crash` (another curve) hits it
        void split_segment(const Msgmt& target, const Msgmt& crash, double ratio) {
            for (size_t k = 0; k < sgmts.size(); k++) {
                if (!(sgmts[k] == target)) continue;
                Msgmt& s = sgmts[k];
                const bool ccw = cross(crash.diff(), s.diff()) > 0;
                const complex uv = lerp(s.fr.uv, s.to.uv, ratio);
                Msgmt tail{this, s.face, Mvert{uv, crash.curv, ccw ? HitL : HitR}, s.to};
                s.to = Mvert{uv, crash.curv, ccw ? HitR : HitL};
                sgmts.insert(sgmts.begin() + std::ptrdiff_t(k) + 1, tail);
                return;
            }
        }
        void post_process() {
            for (size_t i = 0; i < sgmts.size(); i++) sgmts[i].id = int(i);
            for (size_t i = 0; i + 1 < sgmts.size(); i++) {
                sgmts[i].next_id = int(i) + 1;
                sgmts[i + 1].prev_id = int(i);
            }
        }
    };

    class MotorcycleGraph {
    public:
        const VecXc& cfn;
        std::vector<Mport> mports;
        std::vector<Mcurv> mcurvs;

        MotorcycleGraph(const Hmesh& mesh, const VecXc& cfn_, const VecXi& matching, const VecXi& singular)
            : cfn(cfn_), m(mesh), match(matching) {
            for (int h = 0; h < mesh.nH; h++)
                if (mesh.face[h] < 0) throw std::runtime_error("motorcycle graph: closed meshes only");
            gen_ports(singular);
            mcurvs.reserve(mports.size());
            for (const Mport& p : mports) mcurvs.emplace_back(p);
            for (size_t i = 0; i < mcurvs.size(); i++) grow_init(i);
            const int maxSweeps = 4 * mesh.nF + 16;
            for (int sweep = 0;; sweep++) {
                bool all = true;
                for (const Mcurv& c : mcurvs) all = all && c.cache.intersected;
                if (all) break;
                if (sweep >= maxSweeps) throw std::runtime_error("motorcycle graph: a curve never crashed");
                for (size_t i = 0; i < mcurvs.size(); i++) grow_next(i);
            }
            for (Mcurv& c : mcurvs) c.post_process();
        }

    private:
        const Hmesh& m;
        const VecXi& match;

        static bool isccw(complex a, complex b, complex c) { return cross(b - a, c - a) > 0; }
        static bool points_into(complex d, complex a, complex b, complex c) {
            return isccw(a, b, a + d) && isccw(a, a + d, c);
        }

        void gen_ports(const VecXi& singular) {
            int counter = 0;
            for (int v = 0; v < m.nV; v++) {
                if (singular[v] == 0) continue;
                std::vector<Mport> ports;
                m.forVertCcw(m.vert2half[v], [&](int h) {
                    const complex a = cfn(m.crnr[m.next[h]]), b = cfn(m.crnr[m.prev[h]]), c = cfn(m.crnr[h]);
                    if (cross(b - a, c - a) < 0) throw std::invalid_argument("Orientation should be ccw order");
                    int r = 0;
                    for (; r < 4; r++) if (!points_into(get_quater_rot(r), a, b, c) && r > 0) break;
                    std::vector<Mport> temps;
                    for (int i = 0; i < 4; i++) {
                        const complex d = get_quater_rot(r - i);
                        if (points_into(d, a, b, c) || std::arg(d) == std::arg(b - a) || std::arg(d) == std::arg(c - a))
                            temps.emplace_back(-1, Vert{{v, &m}}, Face{{m.face[h], &m}}, a, d);
                    }
                    const complex e = b - a;
                    std::stable_sort(temps.begin(), temps.end(),
                                     [&](const Mport& p0, const Mport& p1) { return dot(p0.dir, e) > dot(p1.dir, e); });
                    for (Mport& p : temps) p.id = counter++;
                    ports.insert(ports.end(), temps.begin(), temps.end());
                    return true;
                });
                const int s = int(ports.size());
                for (int i = 0; i < s; i++) {
                    ports[size_t(i)].prev = ports[size_t((i - 1 + s) % s)].id;
                    ports[size_t(i)].next = ports[size_t((i + 1 + s) % s)].id;
                }
                mports.insert(mports.end(), ports.begin(), ports.end());
            }
        }

        void grow_init(size_t i) {
            const Mport& p = mcurvs[i].port;
            int h = -1;
            const int h0 = m.face2half[p.face.id];
            for (const int hh : {h0, m.next[h0], m.next[m.next[h0]]})
                if (m.head[hh] != p.vert.id && m.tail[hh] != p.vert.id) h = hh;
            add_segment(i, p.dir, p.uv, h);
        }

        void grow_next(size_t i) {
            Mcurv& c = mcurvs[i];
            if (c.cache.intersected) return;
            const int h = m.twin[c.cache.half];
            const int k = (m.edge2half[m.edge[h]] == h ? 1 : -1) * match[m.edge[h]];
            const complex uv0 = lerp(cfn(m.crnr[m.next[h]]), cfn(m.crnr[m.prev[h]]), c.cache.ratio);
            const complex dir = std::polar(1., PI / 2 * k) * c.cache.dir;
            add_segment(i, dir, uv0, opposite_half(uv0, dir, h));
        }

        // the other halfedge of fr's face that the line uv0 + s dir crosses
        int opposite_half(complex origin, complex dir, int fr) const {
            const int h0 = m.face2half[m.face[fr]];
            for (const int h : {h0, m.next[h0], m.next[m.next[h0]]}) {
                if (h == fr) continue;
                const complex a = cfn(m.crnr[m.prev[h]]) - origin, b = cfn(m.crnr[m.next[h]]) - origin;
                if (cross(a, dir) * cross(b, dir) < 0) return h;
            }
            throw std::invalid_argument("motorcycle graph: no opposite half");
        }

        void add_segment(size_t i, complex dir, complex uv0, int h) {
            Mcurv& c = mcurvs[i];
            const complex uv1 = cfn(m.crnr[m.prev[h]]), uv2 = cfn(m.crnr[m.next[h]]);
            double r_ab, r_cd;
            if (!find_extended_intersection(uv0, uv0 + dir, uv1, uv2, r_ab, r_cd))
                throw std::runtime_error("motorcycle graph: curve parallel to the exit edge");
            const complex uv3 = lerp(uv1, uv2, r_cd);
            const int f = m.face[h];
            // nearest strict crossing with a segment of another curve in this face
            double best = 0, bestcd = 0;
            const Msgmt* hit = nullptr;
            for (size_t j = 0; j < mcurvs.size(); j++) {
                if (j == i) continue;
                for (const Msgmt& sg : mcurvs[j].sgmts) {
                    if (sg.face.id != f) continue;
                    double ab, cd;
                    if (find_strict_intersection(uv0, uv3, sg.fr.uv, sg.to.uv, ab, cd) && (!hit || ab < best)) {
                        hit = &sg;
                        best = ab;
                        bestcd = cd;
                    }
                }
            }
            const Face face{{f, &m}};
            if (!hit) {
                c.sgmts.push_back(Msgmt{&c, face, Mvert{uv0}, Mvert{uv3}});
                c.cache = Cache{h, r_cd, dir, false};
                return;
            }
            const Msgmt target = *hit;
            c.sgmts.push_back(Msgmt{&c, face, Mvert{uv0}, Mvert{lerp(uv0, uv3, best), target.curv, HitB}});
            c.cache = Cache{h, r_cd, dir, true};
            target.curv->split_segment(target, c.sgmts.back(), bestcd);
        }
    };

    inline const Msgmt& Msgmt::next() const { return curv->sgmts[size_t(next_id)]; }
    inline const Msgmt& Msgmt::prev() const { return curv->sgmts[size_t(prev_id)]; }
}
#endif
