// metriko (MI355X build) -- QEx faces (reference core/qex/gen_q_face.h:7-39): half 2 i runs
// q-edge i from port1 to port2, half 2 i + 1 back; from the first unvisited half, three
// times take the first half leaving the port before the current end port (prev_id); four
// halves make a quad.
//
// Restated from komietty/metriko include/metriko/core/qex/gen_q_face.h
// (the komietty/metriko authors), which is subject to the terms of the Mozilla Public License v.2.0.
#ifndef METRIKO_QEX_GEN_Q_FACE_H
#define METRIKO_QEX_GEN_Q_FACE_H
#include <unordered_map>
#include "common.h"

namespace metriko::qex {
    inline std::vector<Qface> generate_q_faces(const std::vector<Qport>& qps, const std::vector<Qedge>& qes) {
        std::vector<Qface> qfs;
        std::vector<Qhalf> qhs;
        qhs.reserve(qes.size() * 2);
        for (size_t i = 0; i < qes.size(); i++) {
            qhs.emplace_back(qes[i], int(i * 2 + 0), true);
            qhs.emplace_back(qes[i], int(i * 2 + 1), false);
        }
        std::unordered_map<int, int> first;   // port idx -> first half leaving it
        for (const Qhalf& qh : qhs) first.emplace(qh.port1().idx, qh.idx);
        std::vector<bool> visit(qhs.size(), false);
        size_t start = 0;
        for (;;) {
            while (start < qhs.size() && visit[start]) start++;
            if (start == qhs.size()) break;
            Qface qf;
            qf.qhalfs.push_back(qhs[start]);
            visit[start] = true;
            int port = qhs[start].port2().idx;
            for (int i = 0; i < 3; i++) {
                const auto it = first.find(qps[size_t(qps[size_t(port)].prev_id)].idx);
                if (it == first.end()) throw std::runtime_error("qex: open quad loop");
                qf.qhalfs.push_back(qhs[size_t(it->second)]);
                visit[size_t(it->second)] = true;
                port = qhs[size_t(it->second)].port2().idx;
            }
            if (qf.qhalfs.size() == 4) qfs.push_back(std::move(qf));
        }
        return qfs;
    }
}
#endif
