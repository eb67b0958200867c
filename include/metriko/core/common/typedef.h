// metriko (MI355X build) -- Eigen-free value types for the drop-in headers.
//
// Mirrors the aliases of the reference's core/common/typedef.h:12-40 (MatXd,
// MatXi, VecXd, VecXc, SprsD, SprsC, ...) with small owning containers so the
// headers build without Eigen.  Dense matrices are row-major; element access
// is m(i, j) as in Eigen.
//
// Restated from komietty/metriko include/metriko/core/common/typedef.h
// (Copyright (C) 2025 Saki Komikado), which is subject to the terms of the Mozilla Public License v.2.0.
#ifndef METRIKO_TYPEDEF_H
#define METRIKO_TYPEDEF_H
#include <cassert>
#include <cmath>
#include <complex>
#include <cstdint>
#include <stdexcept>
#include <vector>

namespace metriko {
    constexpr double PI = M_PI;
    constexpr double TwoPI = 2 * M_PI;
    constexpr double EPS = 1e-9;
    using complex = std::complex<double>;

    template<typename T>
    struct MatX {
        int nr = 0, nc = 0;
        std::vector<T> d;
        MatX() = default;
        MatX(int r, int c) : nr(r), nc(c), d(size_t(r) * size_t(c), T()) {}
        static MatX Constant(int r, int c, T v) { MatX m(r, c); std::fill(m.d.begin(), m.d.end(), v); return m; }
        static MatX Zero(int r, int c) { return Constant(r, c, T()); }
        void resize(int r, int c) { nr = r; nc = c; d.assign(size_t(r) * size_t(c), T()); }
        int rows() const { return nr; }
        int cols() const { return nc; }
        T& operator()(int i, int j) { return d[size_t(i) * nc + j]; }
        const T& operator()(int i, int j) const { return d[size_t(i) * nc + j]; }
        T* data() { return d.data(); }
        const T* data() const { return d.data(); }
        T* row(int i) { return d.data() + size_t(i) * nc; }
        const T* row(int i) const { return d.data() + size_t(i) * nc; }
    };

    template<typename T>
    struct VecX {
        std::vector<T> d;
        VecX() = default;
        explicit VecX(int n) : d(size_t(n), T()) {}
        static VecX Constant(int n, T v) { VecX x(n); std::fill(x.d.begin(), x.d.end(), v); return x; }
        static VecX Zero(int n) { return Constant(n, T()); }
        static VecX Ones(int n) { return Constant(n, T(1)); }
        int size() const { return int(d.size()); }
        int rows() const { return int(d.size()); }
        void resize(int n) { d.resize(size_t(n)); }
        void setConstant(T v) { std::fill(d.begin(), d.end(), v); }
        void setZero() { setConstant(T()); }
        T& operator()(int i) { return d[size_t(i)]; }
        const T& operator()(int i) const { return d[size_t(i)]; }
        T& operator[](int i) { return d[size_t(i)]; }
        const T& operator[](int i) const { return d[size_t(i)]; }
        T* data() { return d.data(); }
        const T* data() const { return d.data(); }
        bool empty() const { return d.empty(); }
    };

    struct Row3d {
        double v[3] = {0, 0, 0};
        Row3d() = default;
        Row3d(double x, double y, double z) { v[0] = x; v[1] = y; v[2] = z; }
        double x() const { return v[0]; }
        double y() const { return v[1]; }
        double z() const { return v[2]; }
        double& operator[](int i) { return v[i]; }
        double operator[](int i) const { return v[i]; }
        Row3d operator+(const Row3d& o) const { return {v[0] + o.v[0], v[1] + o.v[1], v[2] + o.v[2]}; }
        Row3d operator-(const Row3d& o) const { return {v[0] - o.v[0], v[1] - o.v[1], v[2] - o.v[2]}; }
        Row3d operator*(double s) const { return {v[0] * s, v[1] * s, v[2] * s}; }
        Row3d operator/(double s) const { return {v[0] / s, v[1] / s, v[2] / s}; }
        double dot(const Row3d& o) const { return v[0] * o.v[0] + v[1] * o.v[1] + v[2] * o.v[2]; }
        Row3d cross(const Row3d& o) const {
            return {v[1] * o.v[2] - v[2] * o.v[1], v[2] * o.v[0] - v[0] * o.v[2], v[0] * o.v[1] - v[1] * o.v[0]};
        }
        double squaredNorm() const { return dot(*this); }
        double norm() const { return std::sqrt(squaredNorm()); }
        Row3d normalized() const { double n = squaredNorm(); return n > 0 ? (*this) / std::sqrt(n) : *this; }
    };

    using MatXd = MatX<double>;
    using MatXi = MatX<int>;
    using MatXc = MatX<complex>;
    using VecXd = VecX<double>;
    using VecXi = VecX<int>;
    using VecXc = VecX<complex>;

    inline Row3d row3(const MatXd& m, int i) { const double* r = m.row(i); return {r[0], r[1], r[2]}; }
    inline void set_row3(MatXd& m, int i, const Row3d& x) { double* r = m.row(i); r[0] = x[0]; r[1] = x[1]; r[2] = x[2]; }

    // C std::round: halves away from zero.
    inline double cround(double x) { return std::round(x); }
}

#endif
