// metriko (MI355X build) -- QGP quantization of the T-mesh edge lengths
// (reference core/quantization/quantization.h compute_quantization / validate_quantization).
// Host code on the T-mesh of the GPU parameterization.  X (integer length per T-edge) starts
// at 0 and grows by generators (loops crossing T-quads side to side, so opposite sides stay
// equal) picked for the lightest zero-length T-edge until X is valid; then up to ten sweeps
// of +-generator moves ordered by compute_weight, each kept when it lowers
// |X ./ R - 1| and stays valid; then one pass over a second basis (shortest loops first).
// |.| is the Euclidean norm summed in index order.
//
// Restated from komietty/metriko include/metriko/core/quantization/quantization.h
// (Copyright (C) 2025 Saki Komikado), which is subject to the terms of the Mozilla Public License v.2.0.
#ifndef METRIKO_QUANTIZATION_H
#define METRIKO_QUANTIZATION_H
#include <cmath>
#include <stdexcept>
#include <tuple>
#include "quantization_basisloop.h"
#include "quantization_weight.h"
#include "quantization_constraint.h"
#include "quantization_evaluation.h"
#include "quantization_validation.h"
namespace metriko {
    inline bool validate_quantization(const Tmesh& tmesh, const VecXd& X) {
        for (const Tquad& tq : tmesh.tquads)
            for (int i = 0; i < 2; i++) {
                int a = 0, b = 0;
                for (const int t : tq.thids_by_side(i)) a += int(X[tmesh.thalfs[size_t(t)].edge().id]);
                for (const int t : tq.thids_by_side(i + 2)) b += int(X[tmesh.thalfs[size_t(t)].edge().id]);
                if (a != b) return false;
            }
        return true;
    }

    namespace detail {
        inline double rel_err(const VecXd& X, const VecXd& R) {
            double s = 0;
            for (int i = 0; i < X.size(); i++) { const double d = X[i] / R[i] - 1.0; s += d * d; }
            return std::sqrt(s);
        }
        inline VecXd add_col(const VecXd& X, const MatXd& G, int c, double sgn) {
            VecXd Y = X;
            for (int i = 0; i < X.size(); i++) Y[i] += sgn * G(i, c);
            return Y;
        }
        inline bool nonneg(const VecXd& X) {
            for (int i = 0; i < X.size(); i++) if (!(X[i] >= 0.)) return false;
            return true;
        }
    }

    inline VecXd compute_quantization(const Tmesh& tmesh, const VecXd& R) {
        const int nE = int(tmesh.nTE);
        VecXd X = VecXd::Zero(nE);
        const MatXd G = construct_generating_vectors(tmesh, R, [](const Comparator& c1, const Comparator& c2) {
            return c1.length / std::max(c1.weight, 1e-9) < c2.length / std::max(c2.weight, 1e-9);
        });
        // first step: cover every zero-length T-edge
        auto any_zero = [&]() { for (int i = 0; i < nE; i++) if (X[i] == 0) return true; return false; };
        while (any_zero()) {
            double mn = 1e+9;
            int thid = 0;
            for (const Thalf& th : tmesh.thalfs) {
                const int te = th.edge().id;
                if (X[te] != 0) continue;
                const double w = compute_weight(R[te], X[te], nE);
                if (w < mn) { mn = w; thid = th.id; }
            }
            const int te = tmesh.thalfs[size_t(thid)].edge().id;
            int col = -1;
            for (int c = 0; c < G.cols() && col < 0; c++) if (G(te, c) != 0) col = c;
            if (col < 0) throw std::runtime_error("quantization: a T-edge no generator covers");
            X = detail::add_col(X, G, col, 1.0);
            if (compute_validation(tmesh, X)) break;
        }
        // second step: weighted +-generator moves
        double e = detail::rel_err(X, R);
        for (int counter = 0; counter < 10;) {
            const double prev_e = e;
            std::vector<std::tuple<double, int, bool>> es;
            for (int j = 0; j < nE; j++) {
                es.emplace_back(compute_weight(R[j], X[j], nE, false), j, false);
                es.emplace_back(compute_weight(R[j], X[j], nE, true), j, true);
            }
            // ties keep their (edge, sign) order (the reference's std::sort leaves it unspecified)
            std::stable_sort(es.begin(), es.end(),
                             [](const auto& a, const auto& b) { return std::get<0>(a) < std::get<0>(b); });
            bool moved = false;
            for (size_t j = 0; j < es.size() && !moved; j++) {
                const int ei = std::get<1>(es[j]);
                for (int c = 0; c < G.cols(); c++) {
                    if (G(ei, c) == 0) continue;
                    const VecXd x1 = detail::add_col(X, G, c, 1.0), x2 = detail::add_col(X, G, c, -1.0);
                    const double n1 = detail::rel_err(x1, R), n2 = detail::rel_err(x2, R);
                    if (n1 <= e && compute_validation(tmesh, x1)) { X = x1; e = n1; moved = true; break; }
                    if (n2 <= e && compute_validation(tmesh, x2)) { X = x2; e = n2; moved = true; break; }
                }
            }
            counter = prev_e == e ? counter + 1 : 0;
        }
        // third step: the shortest-loop basis
        const MatXd G2 = construct_generating_vectors(tmesh, R, [](const Comparator& c1, const Comparator& c2) {
            return c1.length < c2.length;
        });
        for (const Thalf& th : tmesh.thalfs)
            for (int c = 0; c < G2.cols(); c++) {
                if (G2(th.edge().id, c) == 0) continue;
                const VecXd x1 = detail::add_col(X, G2, c, 1.0), x2 = detail::add_col(X, G2, c, -1.0);
                const double n1 = detail::rel_err(x1, R), n2 = detail::rel_err(x2, R);
                if (n1 <= e && detail::nonneg(x1) && compute_validation(tmesh, x1)) { X = x1; e = n1; }
                if (n2 <= e && detail::nonneg(x2) && compute_validation(tmesh, x2)) { X = x2; e = n2; }
            }
        return X;
    }
}
#endif
