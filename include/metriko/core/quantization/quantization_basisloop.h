// metriko (MI355X build) -- QGP quantization: the basis loop of a T-half-edge
// (reference core/quantization/quantization_basisloop.h).  From bgn, a best-first search
// (std::priority_queue ordered by `compare`) crosses T-quads from a side to the opposite
// side (over the twin of every half-edge there) until it returns to bgn; the loop is read
// back through the predecessor map.
//
// Restated from komietty/metriko include/metriko/core/quantization/quantization_basisloop.h
// (Copyright (C) 2025 Saki Komikado), which is subject to the terms of the Mozilla Public License v.2.0.
#ifndef METRIKO_QUANTIZATION_BASISLOOP_H
#define METRIKO_QUANTIZATION_BASISLOOP_H
#include <algorithm>
#include <queue>
#include <stdexcept>
#include <unordered_map>
#include <vector>
#include "metriko/core/tmesh/tmesh.h"
namespace metriko {
    class Comparator {
    public:
        int length;
        double weight;
        int thid;
        bool operator==(const Comparator& r) const { return length == r.length && weight == r.weight; }
        bool operator!=(const Comparator& r) const { return !(*this == r); }
        bool operator<(const Comparator& r) const { return length < r.length || (length == r.length && weight > r.weight); }
        bool operator>(const Comparator& r) const { return length > r.length || (length == r.length && weight < r.weight); }
        bool operator<=(const Comparator& r) const { return !(*this > r); }
        bool operator>=(const Comparator& r) const { return !(*this < r); }
    };

    template <typename Func>
    std::vector<int> gen_basis_loop(const std::vector<Tquad>& tquads, const std::vector<Thalf>& thalfs,
                                    const VecXi& thalf_tquad_table, const VecXi& thalf_sides_table, const VecXd& R,
                                    const Thalf& bgn, Func compare) {
        std::priority_queue<Comparator, std::vector<Comparator>, Func> q(compare);
        std::unordered_map<int, int> m;   // half-edge -> predecessor
        std::vector<int> visited;
        int counter = 0;
        q.push(Comparator{0, 0, bgn.id});
        do {
            const int len = q.top().length;
            const int prev = q.top().thid;
            q.pop();
            const Tquad& tq = tquads[size_t(thalf_tquad_table[prev])];
            const int side = thalf_sides_table[prev];
            std::vector<int> pairs;
            for (size_t i = 0; i < tq.thids.size(); i++)
                if (tq.sides[i] == (side + 2) % 4) pairs.push_back(tq.thids[i]);
            for (const int p : pairs) {
                const int curr = thalfs[size_t(p)].twid;
                if (curr == bgn.id && counter > 0) {
                    std::vector<int> res{prev};
                    int idx = prev;
                    do {
                        idx = m[idx];
                        res.push_back(idx);
                    } while (idx != bgn.id);
                    return res;
                }
                bool on_path = false;
                for (int t = prev; t != bgn.id; t = m[t])
                    if (t == curr) { on_path = true; break; }
                if (!on_path && std::find(visited.begin(), visited.end(), curr) == visited.end()) {
                    q.push(Comparator{len - 1, R[thalfs[size_t(curr)].teid], curr});
                    m.emplace(curr, prev);
                    visited.push_back(curr);
                }
            }
            counter++;
        } while (!q.empty());
        throw std::runtime_error("gen_basis_loop: no loop back to the start");
    }
}
#endif
