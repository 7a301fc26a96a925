// Oracle (test infrastructure, not product code): forward-mode dual numbers in
// double precision. Instantiating the templated FP64 restatement with
// Dual<N> gives the exact derivatives of its output with respect to N seeded
// inputs (the chain rule applied operation by operation), which is the
// reference the hand-written adjoint kernels are checked against. The dual
// arithmetic itself is pinned by FP64 central finite differences of the
// double instantiation (tests/test_oracle_adjoint.py).
//
// Branches and comparisons act on the value part, so the derivative is that
// of the branch taken (clamp, max and threshold tests are piecewise smooth).
#pragma once
#include <algorithm>
#include <cmath>

namespace oracle {

template <int N>
struct Dual {
    static constexpr int n = N;
    double v;
    double d[N];

    Dual() : v(0.0) { zero(); }
    Dual(double x) : v(x) { zero(); }
    void zero() {
        for (int i = 0; i < N; ++i) d[i] = 0.0;
    }
    explicit operator double() const { return v; }
    explicit operator int() const { return (int)v; }

    Dual& operator+=(const Dual& o) { return *this = *this + o; }
    Dual& operator-=(const Dual& o) { return *this = *this - o; }
    Dual& operator*=(const Dual& o) { return *this = *this * o; }
    Dual& operator/=(const Dual& o) { return *this = *this / o; }

    friend Dual operator+(const Dual& a, const Dual& b) {
        Dual r(a.v + b.v);
        for (int i = 0; i < N; ++i) r.d[i] = a.d[i] + b.d[i];
        return r;
    }
    friend Dual operator-(const Dual& a, const Dual& b) {
        Dual r(a.v - b.v);
        for (int i = 0; i < N; ++i) r.d[i] = a.d[i] - b.d[i];
        return r;
    }
    friend Dual operator-(const Dual& a) {
        Dual r(-a.v);
        for (int i = 0; i < N; ++i) r.d[i] = -a.d[i];
        return r;
    }
    friend Dual operator*(const Dual& a, const Dual& b) {
        Dual r(a.v * b.v);
        for (int i = 0; i < N; ++i) r.d[i] = a.d[i] * b.v + a.v * b.d[i];
        return r;
    }
    friend Dual operator/(const Dual& a, const Dual& b) {
        const double q = a.v / b.v;
        Dual r(q);
        for (int i = 0; i < N; ++i) r.d[i] = (a.d[i] - q * b.d[i]) / b.v;
        return r;
    }
    friend bool operator<(const Dual& a, const Dual& b) { return a.v < b.v; }
    friend bool operator>(const Dual& a, const Dual& b) { return a.v > b.v; }
    friend bool operator<=(const Dual& a, const Dual& b) { return a.v <= b.v; }
    friend bool operator>=(const Dual& a, const Dual& b) { return a.v >= b.v; }
    friend bool operator==(const Dual& a, const Dual& b) { return a.v == b.v; }
    friend bool operator!=(const Dual& a, const Dual& b) { return a.v != b.v; }

    // f(a) with f'(a) = df
    static Dual chain(const Dual& a, double f, double df) {
        Dual r(f);
        for (int i = 0; i < N; ++i) r.d[i] = df * a.d[i];
        return r;
    }
    friend Dual exp(const Dual& a) {
        const double e = std::exp(a.v);
        return chain(a, e, e);
    }
    friend Dual log(const Dual& a) { return chain(a, std::log(a.v), 1.0 / a.v); }
    friend Dual sqrt(const Dual& a) {
        const double s = std::sqrt(a.v);
        return chain(a, s, s > 0.0 ? 0.5 / s : 0.0);
    }
    friend Dual sin(const Dual& a) { return chain(a, std::sin(a.v), std::cos(a.v)); }
    friend Dual cos(const Dual& a) { return chain(a, std::cos(a.v), -std::sin(a.v)); }
    friend Dual acos(const Dual& a) {
        const double s = std::sqrt(std::max(1.0 - a.v * a.v, 0.0));
        return chain(a, std::acos(a.v), s > 0.0 ? -1.0 / s : 0.0);
    }
    friend Dual asin(const Dual& a) {
        const double s = std::sqrt(std::max(1.0 - a.v * a.v, 0.0));
        return chain(a, std::asin(a.v), s > 0.0 ? 1.0 / s : 0.0);
    }
    friend Dual floor(const Dual& a) { return Dual(std::floor(a.v)); }
    friend Dual fabs(const Dual& a) { return a.v < 0.0 ? -a : a; }
    friend Dual atan2(const Dual& y, const Dual& x) {
        const double den = x.v * x.v + y.v * y.v;
        Dual r(std::atan2(y.v, x.v));
        if (den > 0.0)
            for (int i = 0; i < N; ++i) r.d[i] = (x.v * y.d[i] - y.v * x.d[i]) / den;
        return r;
    }
    friend Dual pow(const Dual& a, const Dual& b) {
        const double p = std::pow(a.v, b.v);
        Dual r(p);
        const double da = a.v != 0.0 ? b.v * std::pow(a.v, b.v - 1.0) : 0.0;
        const double db = a.v > 0.0 ? p * std::log(a.v) : 0.0;
        for (int i = 0; i < N; ++i) r.d[i] = da * a.d[i] + db * b.d[i];
        return r;
    }
};

template <class R>
struct is_dual {
    static constexpr bool value = false;
};
template <int N>
struct is_dual<Dual<N>> {
    static constexpr bool value = true;
};

// Scalar storage type of the LUT textures an instantiation reads: duals read
// double LUTs (the cloud derivatives hold the atmosphere fixed).
template <class R>
struct LutStore {
    typedef R type;
};
template <int N>
struct LutStore<Dual<N>> {
    typedef double type;
};

inline double value_of(double x) { return x; }
inline double value_of(float x) { return x; }
template <int N>
inline double value_of(const Dual<N>& x) {
    return x.v;
}

}  // namespace oracle
