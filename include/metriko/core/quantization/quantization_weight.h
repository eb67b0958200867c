// metriko (MI355X build) -- QGP quantization: the greedy step weight
// (reference core/quantization/quantization_weight.h).  d = ideal - current (negated for a
// decrease): 1 / (d + 1) once a T-edge is a unit or more short, n / (d + 1) when within one
// unit, n^2 / (1 - d) when already too long; decreasing a zero-length edge is 1e9.
//
// Restated from komietty/metriko include/metriko/core/quantization/quantization_weight.h
// (Copyright (C) 2025 Saki Komikado), which is subject to the terms of the Mozilla Public License v.2.0.
#ifndef METRIKO_QUANTIZATION_WEIGHT_H
#define METRIKO_QUANTIZATION_WEIGHT_H
namespace metriko {
    inline double compute_weight(const double ideal_value, const double curr_value, const int num_tedges_in_tmesh,
                                 const bool minus = false) {
        if (minus && curr_value == 0) return 1e9;
        const double d = (ideal_value - curr_value) * (minus ? -1. : 1.);
        const double n = num_tedges_in_tmesh;
        if (d >= 1.) return 1. / (d + 1.);
        if (d < 0.) return n * n / (1. - d);
        return n / (d + 1.);
    }
}
#endif
