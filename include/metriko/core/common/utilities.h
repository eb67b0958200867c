// metriko (MI355X build) -- small complex / 2D helpers.
//
// Drop-in for the parts of the reference's core/common/utilities.h that the QEx
// extraction uses (get_quater_rot, equal, lerp, dot, cross, get_minmax_int,
// calc_coefficient, find_extended_intersection, find_strict_intersection).  The 2x2
// solves are written as Eigen's 2x2 inverse evaluates them (cofactors times 1/det, then
// the matrix-vector product), so the FP64 results match an Eigen build term by term.
//
// Restated from komietty/metriko include/metriko/core/common/utilities.h
// (Copyright (C) 2025 Saki Komikado), which is subject to the terms of the Mozilla Public License v.2.0.
#ifndef METRIKO_COMMON_UTILITIES_H
#define METRIKO_COMMON_UTILITIES_H
#include <algorithm>
#include <cmath>
#include <stdexcept>
#include <vector>
#include "typedef.h"

namespace metriko {
    inline complex get_quater_rot(int i) {
        switch ((i % 4 + 4) % 4) {
            case 0: return {1, 0};
            case 1: return {0, 1};
            case 2: return {-1, 0};
            default: return {0, -1};
        }
    }
    inline bool equal(const complex a, const complex b, const double tolerance = 1e-10) {
        return std::abs(a - b) < tolerance;
    }
    inline complex lerp(const complex a, const complex b, const double t) { return a * (1. - t) + b * t; }
    inline double dot(const complex a, const complex b) { return a.real() * b.real() + a.imag() * b.imag(); }
    inline double cross(const complex a, const complex b) { return a.real() * b.imag() - a.imag() * b.real(); }

    struct minmax_int { int min_x, min_y, max_x, max_y; };
    inline minmax_int get_minmax_int(const std::vector<complex>& uvs) {
        double x0 = uvs[0].real(), x1 = x0, y0 = uvs[0].imag(), y1 = y0;
        for (const complex& u : uvs) {
            x0 = std::min(x0, u.real()); x1 = std::max(x1, u.real());
            y0 = std::min(y0, u.imag()); y1 = std::max(y1, u.imag());
        }
        return {int(std::floor(x0)), int(std::floor(y0)), int(std::ceil(x1)), int(std::ceil(y1))};
    }

    namespace detail {
        // [m00 m01; m10 m11]^-1 * (x0, x1), Eigen's 2x2 inverse order of operations
        inline void solve2(double m00, double m01, double m10, double m11, double x0, double x1, double& r0,
                           double& r1) {
            const double inv = 1.0 / (m00 * m11 - m10 * m01);
            const double i00 = m11 * inv, i01 = -m01 * inv, i10 = -m10 * inv, i11 = m00 * inv;
            r0 = i00 * x0 + i01 * x1;
            r1 = i10 * x0 + i11 * x1;
        }
    }

    // coefficients (a, b) with target = origin + a (point1 - origin) + b (point2 - origin)
    inline complex calc_coefficient(const complex origin, const complex point1, const complex point2,
                                    const complex target) {
        const complex v1 = point1 - origin, v2 = point2 - origin, d = target - origin;
        double a, b;
        detail::solve2(v1.real(), v2.real(), v1.imag(), v2.imag(), d.real(), d.imag(), a, b);
        return {a, b};
    }

    // intersection of the lines a-b and c-d; false when parallel
    inline bool find_extended_intersection(const complex& a, const complex& b, const complex& c, const complex& d,
                                           double& ratio_a2b, double& ratio_c2d) {
        const double m00 = d.real() - c.real(), m01 = -(b - a).real();
        const double m10 = d.imag() - c.imag(), m11 = -(b - a).imag();
        if (m00 * m11 - m10 * m01 == 0) return false;
        double vx, vy;
        detail::solve2(m00, m01, m10, m11, -c.real() + a.real(), -c.imag() + a.imag(), vx, vy);
        ratio_a2b = vy;
        ratio_c2d = vx;
        return true;
    }
    inline bool find_strict_intersection(const complex& a, const complex& b, const complex& c, const complex& d,
                                         double& ratio_a2b, double& ratio_c2d) {
        return find_extended_intersection(a, b, c, d, ratio_a2b, ratio_c2d) && ratio_a2b > 0 && ratio_a2b < 1 &&
               ratio_c2d > 0 && ratio_c2d < 1;
    }
}
#endif
