// metriko (MI355X build) -- QGP T-mesh from the motorcycle graph.
//
// Drop-in for the reference's core/tmesh/tmesh.h (Tmesh, Tquad, Thalf, Tedge):
//  * T-edges: every curve, in order, is cut after each segment that ends in a hit vertex;
//    T-edge s runs from the first segment after the previous cut to that segment, with
//    half-edges 2 s (canonical, along the curve) and 2 s + 1;
//  * T-quads: from the first unvisited half-edge, choose_next_thalf until the start comes
//    back (a turn at a crash, at a split seen from its right side or at a singular vertex
//    starts the next side), then rotated so the list does not begin inside a side;
//  * th2quad / th2side / th2iter per half-edge and th2sing (the half-edge that leaves each
//    singular vertex).
// Segments are matched as Msgmt::operator== does (face and both end points).  A quad walk
// longer than 4 T-edges + 8 throws (the reference would loop).
//
// Restated from komietty/metriko include/metriko/core/tmesh/tmesh.h
// (Copyright (C) 2025 Saki Komikado), which is subject to the terms of the Mozilla Public License v.2.0.
#ifndef METRIKO_TMESH_H
#define METRIKO_TMESH_H
#include <algorithm>
#include <stdexcept>
#include <utility>
#include <vector>
#include "motorcycle.h"

namespace metriko {
    class Tmesh;

    class Tedge {
    public:
        const int id;
        const Tmesh* tmesh;
        const Msgmt seg_fr;
        const Msgmt seg_to;
        Tedge(const Tmesh* t, int id_, const Msgmt& fr, const Msgmt& to) : id(id_), tmesh(t), seg_fr(fr), seg_to(to) {}
        complex uv_fr() const { return seg_fr.fr.uv; }
        complex uv_to() const { return seg_to.to.uv; }
        MvertType type_fr() const { return seg_fr.fr.type; }
        MvertType type_to() const { return seg_to.to.type; }
        // the curve's segments of this T-edge
        std::vector<Msgmt> segments() const {
            std::vector<Msgmt> out;
            for (const Msgmt& s : seg_fr.curv->sgmts)
                if (s.id >= seg_fr.id && s.id <= seg_to.id) out.push_back(s);
            return out;
        }
    };

    class Thalf {
    public:
        Tmesh* tmesh;
        int id, teid, twid;
        bool cannonical;
        Thalf(Tmesh* t, int teid_, int id_, int twid_, bool c) : tmesh(t), id(id_), teid(teid_), twid(twid_), cannonical(c) {}
        bool operator==(const Thalf& o) const {
            return id == o.id && twid == o.twid && teid == o.teid && cannonical == o.cannonical;
        }
        bool operator!=(const Thalf& o) const { return !(*this == o); }
        const Thalf& twin() const;
        const Thalf& next() const;
        const Thalf& prev() const;
        const Tedge& edge() const;
        complex uv_fr() const;
        complex uv_to() const;
        Msgmt sg_fr() const;
        Msgmt sg_to() const;
        complex dif_fr() const { return sg_fr().diff() * (cannonical ? 1. : -1.); }
        complex dif_to() const { return sg_to().diff() * (cannonical ? 1. : -1.); }
        std::vector<Thalf> adj_thalfs() const;
    };

    class Tquad {
    public:
        int id;
        std::vector<int> thids;
        std::vector<int> sides;

        Tquad(int id_, const std::vector<Mcurv>& mcurvs, const std::vector<Thalf>& thalfs, const Thalf& bgn) : id(id_) {
            int side = 0;
            int cur = bgn.id;
            const size_t cap = thalfs.size() * 2 + 8;
            do {
                thids.push_back(cur);
                sides.push_back(side);
                const auto nx = choose_next_thalf(mcurvs, thalfs, thalfs[size_t(cur)]);
                if (nx.second) side = (side + 1) % 4;
                cur = nx.first.id;
                if (thids.size() > cap) throw std::runtime_error("T-quad does not close");
            } while (cur != bgn.id);
            if (sides.front() == sides.back()) {   // do not start in the middle of a side
                const int first = sides.front();
                size_t n = 0;
                while (n < sides.size() && sides[n] == first) n++;
                std::rotate(sides.begin(), sides.begin() + std::ptrdiff_t(n), sides.end());
                std::rotate(thids.begin(), thids.begin() + std::ptrdiff_t(n), thids.end());
            }
            if (sides.front() == sides.back()) throw std::runtime_error("T-quad with a single side");
        }

        int find_first_thid(int side) const {
            for (size_t i = 0; i < sides.size(); i++) if (sides[i] == side) return thids[i];
            return -1;
        }
        int find_side(const Thalf& th) const {
            for (size_t i = 0; i < thids.size(); i++) if (thids[i] == th.id) return sides[i];
            return -1;
        }
        std::vector<int> thids_by_side(int side, bool reverse = false) const {
            std::vector<int> r;
            for (size_t i = 0; i < thids.size(); i++) {
                const size_t j = reverse ? thids.size() - i - 1 : i;
                if (sides[j] == side) r.push_back(thids[j]);
            }
            return r;
        }

        static std::pair<Thalf, bool> choose_next_thalf(const std::vector<Mcurv>& mcurvs,
                                                        const std::vector<Thalf>& thalfs, const Thalf& curr) {
            auto find_th = [&thalfs](const Msgmt& ms, bool cannonical) -> const Thalf& {
                for (const Thalf& th : thalfs) {
                    const Tedge& te = th.edge();
                    if (th.cannonical == cannonical && (te.seg_fr == ms || te.seg_to == ms)) return th;
                }
                throw std::runtime_error("thalf not found");
            };
            const Tedge& te = curr.edge();
            if (curr.cannonical) {
                if (te.type_to() == HitB) {
                    const complex uv2 = te.seg_to.to.uv, dif = te.seg_to.diff();
                    for (const Msgmt& ms : te.seg_to.to.crash->sgmts) {
                        if (ms.face.id != te.seg_to.face.id) continue;
                        if (equal(ms.fr.uv, uv2) && cross(dif, ms.diff()) > 0) return {find_th(ms, true), true};
                        if (equal(ms.to.uv, uv2) && cross(dif, -ms.diff()) > 0) return {find_th(ms, false), true};
                    }
                    throw std::runtime_error("thalf not found");
                }
                if (te.type_to() == HitR) return {find_th(te.seg_to.to.crash->sgmts.back(), false), true};
                return {thalfs[size_t(curr.id) + 2], false};
            }
            if (te.seg_fr == te.seg_fr.curv->sgmts.front()) {
                const Mcurv& c = mcurvs[size_t(te.seg_fr.curv->port.prev)];
                return {find_th(c.sgmts.front(), true), true};
            }
            if (te.type_fr() == HitR) return {find_th(te.seg_fr.fr.crash->sgmts.back(), false), true};
            return {thalfs[size_t(curr.id) - 2], false};
        }
    };

    class Tmesh {
    public:
        std::vector<Tquad> tquads;
        std::vector<Thalf> thalfs;
        std::vector<Tedge> tedges;
        VecXi th2sing, th2quad, th2side, th2iter;
        size_t nTQ = 0, nTE = 0, nTH = 0;

        Tmesh() = default;
        Tmesh(const Tmesh&) = delete;
        Tmesh& operator=(const Tmesh&) = delete;
        explicit Tmesh(const std::vector<Mcurv>& mcurvs) {
            for (const Mcurv& mc : mcurvs) {
                Msgmt start = mc.sgmts.front();
                for (const Msgmt& s : mc.sgmts) {
                    if (s.to.type == None) continue;
                    const int id = int(tedges.size());
                    tedges.emplace_back(this, id, start, s);
                    thalfs.emplace_back(this, id, 2 * id, 2 * id + 1, true);
                    thalfs.emplace_back(this, id, 2 * id + 1, 2 * id, false);
                    if (s.next_id != -1) start = mc.sgmts[size_t(s.next_id)];
                }
            }
            std::vector<bool> visit(thalfs.size(), false);
            for (size_t b = 0; b < thalfs.size(); b++) {
                if (visit[b]) continue;
                tquads.emplace_back(int(tquads.size()), mcurvs, thalfs, thalfs[b]);
                for (const int t : tquads.back().thids) visit[size_t(t)] = true;
            }
            nTE = tedges.size();
            nTH = thalfs.size();
            nTQ = tquads.size();
            th2quad.resize(int(nTH));
            th2side.resize(int(nTH));
            th2iter.resize(int(nTH));
            th2sing = VecXi::Constant(int(nTH), -1);
            for (size_t q = 0; q < nTQ; q++)
                for (size_t j = 0; j < tquads[q].thids.size(); j++) {
                    const int t = tquads[q].thids[j];
                    th2quad[t] = int(q);
                    th2side[t] = tquads[q].sides[j];
                    th2iter[t] = int(j);
                }
            for (const Mcurv& mc : mcurvs) {
                const auto it = std::find_if(thalfs.begin(), thalfs.end(),
                                             [&](const Thalf& th) { return th.edge().seg_fr == mc.sgmts.front(); });
                if (it == thalfs.end()) throw std::runtime_error("T-mesh: a curve without a T-edge");
                th2sing[it->id] = mc.port.vert.id;
            }
        }

        int next_thid(int curr) const {
            const Tquad& q = tquads[size_t(th2quad[curr])];
            return q.thids[size_t(th2iter[curr] + 1) % q.thids.size()];
        }
        int prev_thid(int curr) const {
            const Tquad& q = tquads[size_t(th2quad[curr])];
            return q.thids[(size_t(th2iter[curr]) + q.thids.size() - 1) % q.thids.size()];
        }
    };

    inline const Tedge& Thalf::edge() const { return tmesh->tedges[size_t(teid)]; }
    inline const Thalf& Thalf::twin() const { return tmesh->thalfs[size_t(twid)]; }
    inline const Thalf& Thalf::next() const { return tmesh->thalfs[size_t(tmesh->next_thid(id))]; }
    inline const Thalf& Thalf::prev() const { return tmesh->thalfs[size_t(tmesh->prev_thid(id))]; }
    inline complex Thalf::uv_fr() const { return cannonical ? edge().uv_fr() : edge().uv_to(); }
    inline complex Thalf::uv_to() const { return cannonical ? edge().uv_to() : edge().uv_fr(); }
    inline Msgmt Thalf::sg_fr() const { return cannonical ? edge().seg_fr : edge().seg_to; }
    inline Msgmt Thalf::sg_to() const { return cannonical ? edge().seg_to : edge().seg_fr; }
    inline std::vector<Thalf> Thalf::adj_thalfs() const {
        std::vector<Thalf> res;
        const MvertType type = cannonical ? sg_to().to.type : sg_to().fr.type;
        const Thalf& nx = next();
        if (type == HitR) {
            res.push_back(nx);
            res.push_back(nx.twin().next());
        } else if (type == HitB || type == HitL) {
            res.push_back(nx);
            res.push_back(twin().prev().twin());
        }
        return res;
    }
}
#endif
